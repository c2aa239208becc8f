"""The genome-scan workload as specified: genes with their own Yule trees of
12-64 taxa and their own foreground branch, lognormal codon counts (median
400, sigma 0.7, clipped to [50, 5000]), parameters kappa ~ U(1, 4),
w0 ~ U(0.01, 0.5), w2 ~ U(1.5, 6) (scan.simulate_genes at its defaults),
fitted under H1 and H0 on batch engines with FP32 conditional likelihoods
and an FP64 finish (scan.scan_batched(precision="fp32")): the FP32 chains
stop at their rounding-noise floor, the FP64 finish takes each fit to the
FP64 maximum.  The reference runs one codeml process per gene
(codeml.c:357) with H1 and H0 as two control files.

Checks: every fit converges (H_end rule, a vanishing projected gradient or
a vanishing Newton step), except where the maximum lies on the boundary of
the model, at infinity of codeml's logistic variables (a class proportion
going to 0): there x cannot converge, lnL creeps towards its supremum as the
proportion shrinks, and the fit ends by the stall rule or without a descent
step with that proportion small (below 0.05 here); the reported lnL at the
H1 and H0 estimates of a sample of genes equals the oracle's likelihood at
those parameters to 1e-9; the statistic 2 dlnL of those genes equals that
of per-gene FP64 fits on their own engines to 1e-4 (2e-3 where a fit ends
on such a boundary)."""
import numpy as np
import pytest

from cucodeml_amd import engine, scan
from oracle import oracle

pytestmark = pytest.mark.gpu

NGENES = 48


def _natural(x, nb, fix_omega):
    a, b = np.exp(x[nb + 1]), np.exp(x[nb + 2])
    p0, p1 = a / (1 + a + b), b / (1 + a + b)
    return x[nb], p0, p1, x[nb + 3], (1.0 if fix_omega else x[nb + 4])


def test_scan_workload_fp32_with_fp64_finish():
    genes = scan.simulate_genes(NGENES)
    groups = scan.batch_groups(genes, 64)
    print(f"case: scan workload, {NGENES} genes, FP32 + FP64 finish, {len(groups)} batch "
          f"groups, taxa {min(g.tree.ns for g in genes)}-{max(g.tree.ns for g in genes)}, "
          f"patterns {min(g.data.npatt for g in genes)}-{max(g.data.npatt for g in genes)}",
          flush=True)
    assert len(groups) >= 2
    res, wall = scan.scan_batched(genes, batch_genes=64, concurrent=4, precision="fp32")
    ev32 = sum(r["fp32_evaluations"] for r in res)
    ev = sum(r["evaluations"] for r in res)
    print(f"  wall {wall:.2f} s, {NGENES / wall * 3600:.0f} genes/hour; evaluations "
          f"{ev32} fp32 + {ev - ev32} fp64 finish", flush=True)
    stops = {}
    for r in res:
        stops[r["stop"]] = stops.get(r["stop"], 0) + 1
    print(f"  stop reasons (H1, H0): {stops}", flush=True)
    conv = ("hend", "pgrad", "newton")
    for i, r in enumerate(res):
        nb = genes[i].tree.nbranch
        for hyp, stop, x in (("H1", r["stop"][0], r["x1"]), ("H0", r["stop"][1], r["x0"])):
            if stop in conv:
                continue
            _, p0, p1, _, _ = _natural(np.array(x), nb, hyp == "H0")
            print(f"  {r['gene']} {hyp} stopped by {stop} at p0 {p0:.3e} p1 {p1:.3e} "
                  f"p2 {1 - p0 - p1:.3e}", flush=True)
            assert stop in ("stall", "nodescent") and min(p0, p1, 1 - p0 - p1) < 0.05, \
                (r["gene"], hyp, stop)
        assert r["lnL1"] >= r["lnL0"] - 1e-6
    # oracle at the estimates of a sample: the smallest, the largest and two
    # in between by pattern count x taxa
    order = sorted(range(NGENES), key=lambda i: genes[i].data.npatt * genes[i].tree.ns)
    sample = [order[0], order[NGENES // 3], order[2 * NGENES // 3], order[-1]]
    for i in sample:
        g, r = genes[i], res[i]
        nb = g.tree.nbranch
        t0 = g.tree.branch.copy()
        for hyp, key, x in (("H1", "lnL1", np.array(r["x1"])), ("H0", "lnL0", np.array(r["x0"]))):
            g.tree.set_branch_vector(x[:nb])
            want, _, _ = oracle.lnl(g.tree, g.data, *_natural(x, nb, hyp == "H0"))
            g.tree.branch[:] = t0
            print(f"  {g.name} ({g.tree.ns} taxa, {g.data.npatt} patterns) {hyp}: lnL "
                  f"{r[key]:.8f} oracle {want:.8f}", flush=True)
            assert abs(r[key] - want) <= 1e-9 * abs(want)
        single = scan.fit_gene(g, precision="fp64")
        print(f"  {g.name}: 2dlnL {r['stat']:.6f}, per-gene FP64 fits {single['stat']:.6f} "
              f"(lnL1 {r['lnL1']:.6f} / {single['lnL1']:.6f}, lnL0 {r['lnL0']:.6f} / "
              f"{single['lnL0']:.6f})", flush=True)
        conv = ("hend", "pgrad", "newton")
        boundary = not all(st in conv for st in tuple(r["stop"]) + tuple(single["stop"]))
        assert abs(r["stat"] - single["stat"]) <= (2e-3 if boundary else 1e-4)
