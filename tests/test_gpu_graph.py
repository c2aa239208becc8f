"""Evaluations replayed from a captured HIP graph (bsa_set_graph) give
bit-for-bit the results of the same engine code launched eagerly: several
replays of both graphs (lnL only, lnL + gradient) at changing parameters,
with eager work on another engine and host synchronisation between them, in
FP64 and FP32; and a fit on a replaying engine equals the eager fit."""
import numpy as np
import pytest

from cucodeml_amd import engine
from . import fixtures

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_eager(precision):
    print(f"case: graph replay cfg2s {precision}", flush=True)
    fx, tree, data = fixtures.load("cfg2s")
    kappa, p0, p1, w0, w2 = fixtures.params(fx)
    x0 = engine.x_from_params(tree.branch_vector(), kappa, p0, p1, w0, w2)
    eg = engine.BranchSiteEngine(tree, data, precision=precision)
    ee = engine.BranchSiteEngine(tree, data, precision=precision)
    eg.use_graph(True)
    try:
        for k in range(6):
            x = x0 * (1 + 1e-3 * np.sin(np.arange(len(x0)) + k))
            if k % 2:
                a, b = eg.lnl(x), ee.lnl(x)
                assert a == b
            else:
                (fa, ga), (fb, gb) = eg.lnl_grad(x), ee.lnl_grad(x)
                assert fa == fb and np.array_equal(ga, gb)
                # per-site values left by the replay are the eager ones too
                assert np.array_equal(eg.fhK()[0], ee.fhK()[0])
        t = eg.timing()
        assert t["sum_pmat"] > 0 and t["sum_prune"] == 0
    finally:
        eg.close()
        ee.close()


def test_fit_with_graph_replay_equals_eager_fit():
    """bsa_fit on a graph-replaying engine evaluates value and gradient
    together at every trial (one graph); on an eager engine a trial runs the
    forward pass alone and the sweep follows only at accepted points.  Both
    are the same minimiser on the same values: same estimates, -lnL,
    iterations and evaluations, bitwise."""
    print("case: fit cfg2s H1 with graph replay vs eager", flush=True)
    from cucodeml_amd import codeml
    fx, tree, data = fixtures.load("cfg2s")
    x0 = codeml.default_start(tree, 2.0, 1.5, False)
    eg = engine.BranchSiteEngine(tree, data)
    ee = engine.BranchSiteEngine(tree, data)
    eg.use_graph(True)
    try:
        lo, hi = ee.default_bounds()
        xg, fg, ig = eg.fit(x0, lo, hi, max_iter=60)
        xe, fe, ie = ee.fit(x0, lo, hi, max_iter=60)
    finally:
        eg.close()
        ee.close()
    print(f"  graph: -lnL {fg:.10f}, {ig['iterations']} iterations, {ig['evaluations']} "
          f"evaluations; eager: -lnL {fe:.10f}, {ie['iterations']} iterations, "
          f"{ie['evaluations']} evaluations", flush=True)
    assert np.array_equal(xg, xe) and fg == fe
    assert ig["iterations"] == ie["iterations"] and ig["evaluations"] == ie["evaluations"]
