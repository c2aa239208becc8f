"""Maximum-likelihood fits with the engine's own optimiser (bsa_fit, in place
of ming2 at codeml.c:688) against the reference CPU codeml's fits of the same
data from the same seeded starts, through the Python API and through the
codeml-compatible driver (control file in, ``out`` and ``lnf`` files out),
with the adjoint gradient, with forward differences, and with FP32
partials; and the likelihood-ratio statistic of H1 over H0.

ml16 (16 taxa x 600 codons, foreground on the longest interior branch) has
an interior H1 maximum (p2 = 0.19, w2 = 5.1): there every estimate (each
branch length, kappa, p0, p1, w0, w2), lnL and 2 dlnL are held to 1e-4
(codeml.c:688, 738-751).  cfg0 (8 taxa x 300 codons) is the documented
exception: its H1 maximum has p0 = p1 = 0, where w2 is not identified, so
its estimates are held only as far as the surface determines them."""
import os
import tempfile

import numpy as np
import pytest

from cucodeml_amd import codeml
from . import fixtures

pytestmark = pytest.mark.gpu

# lnL of a converged fit: the reference prints 6 decimals and stops on
# relative changes of 1e-8 (ming2's e); estimates agree to the precision the
# reference's convergence leaves (its x is printed with 6 decimals).  With
# forward differences (the reference's own gradient) w2 is the least
# determined estimate of cfg0's H1: at its MLE (p0 = p1 = 0) lnL changes by
# less than 1e-5 over dw2 = 0.02, which is what the difference gradient can
# resolve there.
TOL_LNL = 2e-4
TOL_X = {"adjoint": 2e-3, "fd": 5e-2}
# the bar where the maximum is identifiable (ml16).  The reference's own
# estimate of w2 there (SE 2.9) is 7.7e-5 from the exact maximum (Newton
# iterations on the adjoint gradient's differences land on 5.057031, the
# reference printed 5.057108), which is as close as its stopping rule gets.
TOL_ML = 1e-4
# With forward differences (gradientB, the reference's gradient) a fit ends
# where the difference gradient no longer gives a descent direction
# (BSA_STOP_NODESCENT): that is its noise floor, not a failure.  There the
# flat w2 of ml16's H1 is resolved to 5e-4; every other estimate to 1e-4.
TOL_W2_FD = 5e-4


def hyp_args(hyp):
    return (False, 1.5) if hyp == "H1" else (True, 1.0)


def compare(fit, ml, what, tol_x, tol_lnl=TOL_LNL, fd=False, tol_last=None, boundary=False):
    x = fit.x_untransformed()
    ref = np.array(ml["mle_x"])
    dx = np.abs(x - ref)
    print(f"  {what}: lnL {fit.lnL:.6f} vs reference {ml['mle_lnL']:.6f} "
          f"(d {fit.lnL - ml['mle_lnL']:+.2e}); max |dx| {dx.max():.2e} at {int(dx.argmax())}; "
          f"{fit.iterations} iterations, {fit.evaluations} evaluations, stop {fit.stop}, "
          f"{fit.seconds:.2f} s", flush=True)
    # boundary: cfg0's H1 maximum has p0 = p1 = 0, at infinity of the
    # logistic variables, where x cannot converge and a fit ends by the stall
    # rule or without a descent step (lnL is held all the same)
    assert fit.converged or (fd and fit.stop == "nodescent") or \
        (boundary and fit.stop in ("stall", "nodescent")), fit.stop
    # the fit may not end below the reference's optimum, and should not
    # improve on it by more than the reference's own convergence slack
    assert abs(fit.lnL - ml["mle_lnL"]) <= tol_lnl
    if tol_last is None:
        assert dx.max() <= tol_x
    else:
        assert dx[:-1].max() <= tol_x and dx[-1] <= tol_last


@pytest.mark.parametrize("gradient", ["adjoint", "fd"])
@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_fit_matches_reference_mle(hyp, gradient):
    print(f"case: fit cfg0 {hyp} {gradient}", flush=True)
    fx, tree, data = fixtures.load("cfg0")
    ml = fx["ml"][hyp]
    fo, om = hyp_args(hyp)
    fit = codeml.fit_model_a(tree, data, fo, om, start=ml["start"], gradient=gradient)
    compare(fit, ml, f"{hyp} {gradient}", TOL_X[gradient], fd=gradient == "fd",
            boundary=hyp == "H1")


def test_fp32_fit_with_fp64_polish_reaches_the_reference_optimum():
    print("case: fit cfg0 H1 fp32 + fp64 polish", flush=True)
    fx, tree, data = fixtures.load("cfg0")
    ml = fx["ml"]["H1"]
    fit = codeml.fit_model_a(tree, data, False, 1.5, start=ml["start"], precision="fp32")
    ex = fit.extra
    print(f"  fp32 stage lnL {ex['fp32_lnL']:.6f} ({ex['fp32_iterations']} iterations); "
          f"after fp64 polish {fit.lnL:.6f} ({fit.iterations} iterations in all)", flush=True)
    # the FP32 stage gets close (it stalls at its rounding-noise floor)...
    assert ex["fp32_lnL"] >= ml["mle_lnL"] - 0.05
    # ...and the polished fit is the reference's
    compare(fit, ml, "H1 fp32+polish", TOL_X["adjoint"], boundary=True)


def test_driver_control_file_to_out_file():
    print("case: driver cfg0 H0 from control file", flush=True)
    fx, _, _ = fixtures.load("cfg0")
    ml = fx["ml"]["H0"]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
        open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
        open(os.path.join(d, "in.codeml"), "w").write(" ".join(f"{v:.12f}" for v in ml["start"]))
        open(os.path.join(d, "codeml.ctl"), "w").write(fixtures.CTL.format(fix_omega=1,
                                                                           omega=1.0))
        fit = codeml.run(os.path.join(d, "codeml.ctl"))
        lnl, npar, x = fixtures.parse_out(open(os.path.join(d, "out")).read())
    assert npar == len(ml["mle_x"])
    assert abs(lnl - ml["mle_lnL"]) <= TOL_LNL
    assert np.max(np.abs(np.array(x) - np.array(ml["mle_x"]))) <= TOL_X["adjoint"]
    assert abs(lnl - round(fit.lnL, 6)) < 1e-9


def test_driver_fixed_parameters_match_reference_lnl():
    # in.codeml starting with -1: evaluate only (treesub.c:3925-3927)
    print("case: driver cfg0 fixed parameters", flush=True)
    fx, tree, _ = fixtures.load("cfg0")
    p = fx["fixed_params"]
    vals = list(fx["true_branch"]) + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
        open(os.path.join(d, "tree.nwk"), "w").write(fx["newick"] + "\n")
        open(os.path.join(d, "in.codeml"), "w").write("-1 " + " ".join(f"{v:.12f}" for v in vals))
        open(os.path.join(d, "codeml.ctl"), "w").write(fixtures.CTL.format(fix_omega=0,
                                                                           omega=1.5))
        codeml.run(os.path.join(d, "codeml.ctl"))
        lnl, _, _ = fixtures.parse_out(open(os.path.join(d, "out")).read())
    assert abs(lnl - fx["fixed_lnL"]) <= 2e-6


def test_driver_several_data_sets_equal_one_at_a_time():
    # ndata = 3: the alignments follow one another in the sequence file and
    # share the tree file (codeml.c:315-343); the driver fits them together
    # on one batch engine.  Each data set's section of out must equal a run
    # of the driver on that data set alone.
    print("case: driver ndata = 3 (cfg0, a simulated set, cfg0 rows reversed), H1", flush=True)
    from cucodeml_amd import simulate
    fx, tree, _ = fixtures.load("cfg0")
    rng = np.random.default_rng(31)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs_b, _ = simulate.simulate_model_a(tree, 90, 2.5, 0.2, 3.0, 0.5, 0.3, pi, rng)
    lines = fx["phylip"].strip().splitlines()
    rev = lines[0] + "\n" + "\n".join(reversed([ln for ln in lines[1:] if ln.strip()])) + "\n"
    texts = [fx["phylip"].strip() + "\n",
             f"{len(tree.names)} {len(seqs_b[0])}\n"
             + "".join(f"{n}  {q}\n" for n, q in zip(tree.names, seqs_b)), rev]
    ctl = fixtures.CTL.format(fix_omega=0, omega=1.5)
    single = []
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
        for t in texts:
            open(os.path.join(d, "seq.phy"), "w").write(t)
            open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
            single.append(codeml.run(os.path.join(d, "codeml.ctl")))
        open(os.path.join(d, "seq.phy"), "w").write("\n".join(texts))
        open(os.path.join(d, "codeml.ctl"), "w").write(ctl + "ndata = 3\n")
        fits = codeml.run(os.path.join(d, "codeml.ctl"))
        out = open(os.path.join(d, "out")).read()
        rst = open(os.path.join(d, "rst")).read()
    assert len(fits) == 3
    sections = out.split("\n\nData set ")[1:]
    assert [s.split("\n", 1)[0] for s in sections] == ["1", "2", "3"]
    assert rst.count("Bayes Empirical Bayes (BEB) probabilities") == 3
    for k, (f, g, sec) in enumerate(zip(fits, single, sections)):
        lnl, npar, x = fixtures.parse_out(sec)
        print(f"  data set {k + 1}: lnL {f.lnL:.6f} alone {g.lnL:.6f}; {f.iterations} / "
              f"{g.iterations} iterations", flush=True)
        assert f.stop == g.stop and f.stop != "limit"
        assert abs(f.lnL - g.lnL) <= 1e-6 * abs(g.lnL)
        assert abs(lnl - round(f.lnL, 6)) < 1e-9 and npar == f.np
    # cfg0 with its rows in reverse order is the same problem
    assert abs(fits[0].lnL - fits[2].lnL) <= 1e-6 * abs(fits[0].lnL)


def test_lrt_statistic():
    print("case: LRT cfg0", flush=True)
    fx, tree, data = fixtures.load("cfg0")
    fits = {}
    for hyp in ("H1", "H0"):
        fo, om = hyp_args(hyp)
        fits[hyp] = codeml.fit_model_a(tree, data, fo, om, start=fx["ml"][hyp]["start"])
    r = codeml.lrt(fits["H0"], fits["H1"])
    ref = 2 * (fx["ml"]["H1"]["mle_lnL"] - fx["ml"]["H0"]["mle_lnL"])
    print(f"  2dlnL {r['stat']:.6f} vs reference {ref:.6f}; p {r['p_mixture']:.4f}", flush=True)
    assert abs(r["stat"] - ref) <= 2 * TOL_LNL


@pytest.mark.parametrize("gradient", ["adjoint", "fd"])
@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_fit_matches_reference_mle_interior(hyp, gradient):
    print(f"case: fit ml16 {hyp} {gradient}", flush=True)
    fx, tree, data = fixtures.load("ml16")
    ml = fx["ml"][hyp]
    fo, om = hyp_args(hyp)
    fit = codeml.fit_model_a(tree, data, fo, om, start=ml["start"], gradient=gradient)
    fd = gradient == "fd"
    compare(fit, ml, f"ml16 {hyp} {gradient}", TOL_ML, TOL_ML, fd=fd,
            tol_last=TOL_W2_FD if fd and hyp == "H1" else None)
    if hyp == "H1":
        assert fit.proportions[2] + fit.proportions[3] > 0.1 and fit.w2 > 2


def test_lrt_statistic_interior():
    print("case: LRT ml16", flush=True)
    fx, tree, data = fixtures.load("ml16")
    fits = {}
    for hyp in ("H1", "H0"):
        fo, om = hyp_args(hyp)
        fits[hyp] = codeml.fit_model_a(tree, data, fo, om, start=fx["ml"][hyp]["start"])
    r = codeml.lrt(fits["H0"], fits["H1"])
    ref = 2 * (fx["ml"]["H1"]["mle_lnL"] - fx["ml"]["H0"]["mle_lnL"])
    print(f"  2dlnL {r['stat']:.6f} vs reference {ref:.6f}; p {r['p_mixture']:.3e}", flush=True)
    assert abs(r["stat"] - ref) <= TOL_ML


def test_fp32_fit_with_fp64_polish_interior():
    print("case: fit ml16 H1 fp32 + fp64 polish", flush=True)
    fx, tree, data = fixtures.load("ml16")
    ml = fx["ml"]["H1"]
    fit = codeml.fit_model_a(tree, data, False, 1.5, start=ml["start"], precision="fp32")
    print(f"  fp32 stage lnL {fit.extra['fp32_lnL']:.6f} stop {fit.extra['fp32_stop']}",
          flush=True)
    compare(fit, ml, "ml16 H1 fp32+polish", TOL_ML, TOL_ML)


def _driver_dir(d, fx, ctl, in_codeml=None, newick=None):
    open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
    open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{newick or fx['newick']}\n")
    if in_codeml is not None:
        open(os.path.join(d, "in.codeml"), "w").write(in_codeml)
    open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
    return os.path.join(d, "codeml.ctl")


@pytest.mark.parametrize("hyp", ["H1", "H0"])
def test_driver_fix_blength_2_matches_reference(hyp):
    """fix_blength = 2: branch lengths held at the tree file's values, only
    kappa, p0, p1, w0 (and w2) estimated (treesub.c:3729-3730,
    codeml.c:617-627, 741-751), against the reference's fit of ml16 from the
    same start."""
    print(f"case: driver fix_blength = 2 ml16 {hyp}", flush=True)
    fx, _, _ = fixtures.load("ml16")
    ml = fx["ml_fix_blength2"][hyp]
    fo, om = (0, 1.5) if hyp == "H1" else (1, 1.0)
    ctl = fixtures.CTL.format(fix_omega=fo, omega=om).replace("fix_blength = 0",
                                                              "fix_blength = 2")
    with tempfile.TemporaryDirectory() as d:
        fit = codeml.run(_driver_dir(d, fx, ctl, " ".join(f"{v:.12f}" for v in ml["start"])))
        text = open(os.path.join(d, "out")).read()
    lnl, npar, x = fixtures.parse_out(text)
    print(f"  lnL {lnl:.6f} vs reference {ml['mle_lnL']:.6f}; x {x} vs {ml['mle_x']}; "
          f"stop {fit.stop}", flush=True)
    assert "lnL(ntime: 0" in text
    assert npar == len(ml["mle_x"])
    assert abs(lnl - ml["mle_lnL"]) <= TOL_ML
    assert np.max(np.abs(np.array(x) - np.array(ml["mle_x"]))) <= TOL_ML
    fx0, tree, _ = fixtures.load("ml16")
    # the tree file carries the lengths with 10 decimals
    assert np.max(np.abs(fit.branch - tree.branch_vector())) < 1e-10


def test_driver_lnf_matches_reference():
    """The lnf file the driver writes for a fixed-parameter run (in.codeml
    starting with -1) against the reference's lnf for the same input:
    header, pattern order and codon columns exactly, log f and f to their
    print precision."""
    print("case: driver lnf ml16", flush=True)
    fx, tree, _ = fixtures.load("ml16")
    p = fx["fixed_params"]
    vals = list(fx["true_branch"]) + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
    ctl = fixtures.CTL.format(fix_omega=0, omega=1.5)
    with tempfile.TemporaryDirectory() as d:
        codeml.run(_driver_dir(d, fx, ctl, "-1 " + " ".join(f"{v:.12f}" for v in vals)))
        got = open(os.path.join(d, "lnf")).read()
    ref = fx["fixed_lnf_text"]
    rg = [r.split() for r in got.splitlines()]
    rr = [r.split() for r in ref.splitlines()]
    assert len(rg) == len(rr) and rg[0] == rr[0]
    for a, b in zip(rg, rr):
        if len(b) <= 5:
            assert a == b
            continue
        assert a[:2] == b[:2] and a[5:] == b[5:]
        assert abs(float(a[2]) - float(b[2])) <= 1e-10 * abs(float(b[2])) + 2e-10
        assert abs(float(a[4]) - float(b[4])) <= 1.5e-4


@pytest.mark.parametrize("gradient", ["adjoint", "fd"])
def test_fit_asks_gradients_only_at_accepted_points(gradient):
    """bsa_fit evaluates a trial point's likelihood alone and asks for the
    gradient only where it accepts the point (the adjoint sweep of the kept
    forward pass, or gradientB's differences).  The trajectory is that of an
    optimiser given value and gradient at every trial point, bitwise: the
    same minimiser on a host objective that evaluates both (lnl_grad or
    gradientB) stops at the same x and -lnL after the same iterations and
    evaluations."""
    from cucodeml_amd import engine
    print(f"case: gradient only at accepted points, ml16 H1, {gradient}", flush=True)
    fx, tree, data = fixtures.load("ml16")
    eng = engine.BranchSiteEngine(tree, data)
    try:
        x0 = codeml.default_start(tree, 2.0, 1.5, False)
        lo, hi = eng.default_bounds()
        xs, fs, ins = eng.fit(x0, lo, hi, max_iter=40, gradient=gradient)
        both = eng.lnl_grad if gradient == "adjoint" else eng.gradientB
        xb, fb, inb = engine.minimize(both, x0, lo, hi, max_iter=40)
    finally:
        eng.close()
    print(f"  fit: -lnL {fs:.10f}, {ins['iterations']} iterations, {ins['evaluations']} "
          f"evaluations; both at every trial: -lnL {fb:.10f}, {inb['iterations']} iterations, "
          f"{inb['evaluations']} evaluations", flush=True)
    assert np.array_equal(xs, xb) and fs == fb
    assert ins["iterations"] == inb["iterations"]
    # bsa_fit counts its final evaluation at the returned x too
    assert ins["evaluations"] == inb["evaluations"] + 1


def test_fp64_finish_on_a_large_tree_takes_preconditioned_newton_steps():
    """A tree of 100 taxa has 202 variables (H1): more than 160, so the
    Newton iterations of the FP64 finish take their steps from conjugate
    gradients preconditioned by the model parameters' exact Hessian block
    (bsa_fit.cpp pcg_step), and codeml.finish_fp64 runs them alone
    (FINISH_NEWTON_LARGE).  From the FP32 fit's estimates they must land on
    the maximum that the FP64 BFGS fit from the same start finds (or above
    it), and say so (stop "newton")."""
    from cucodeml_amd import engine, formats, simulate
    print("case: FP32 fit + FP64 Newton finish (preconditioned steps), 100 taxa x 400 codons",
          flush=True)
    rng = np.random.default_rng(31)
    tree = simulate.yule_tree(100, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 400, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    start = codeml.default_start(tree, 2.0, 1.5, False)
    assert len(start) > codeml.DENSE_NEWTON_MAX
    f64 = codeml.fit_model_a(tree, data, precision="fp64", start=start)
    f32 = codeml.fit_model_a(tree, data, precision="fp32", start=start)
    ex = f32.extra
    print(f"  fp64 BFGS fit: lnL {f64.lnL:.8f} ({f64.iterations} iterations, stop {f64.stop}); "
          f"fp32 fit {ex['fp32_lnL']:.8f} ({ex['fp32_stop']}), after the FP64 Newton finish "
          f"{f32.lnL:.8f} ({f32.iterations} iterations in all, {f32.evaluations} evaluations, "
          f"stop {f32.stop}); w2 {f64.w2:.6f} / {f32.w2:.6f}", flush=True)
    assert f32.stop == "newton" and f32.converged
    assert f32.lnL >= f64.lnL - 1e-6
    assert abs(f32.lnL - f64.lnL) <= 1e-4
    for name in ("kappa", "p0", "p1", "w0"):
        assert abs(getattr(f32, name) - getattr(f64, name)) <= 1e-3, name
