"""Genome scan: branch-site tests of many genes on one GPU.

Each gene is its own tree (with its own ``#1`` branch) and alignment.  A gene
is fitted under H1 and H0 (w2 = 1 fixed, started from the H1 estimates) and
gets the likelihood-ratio statistic; this is the reference's per-gene
Forestry run of codeml (codeml.c:578-854), repeated per gene by separate
processes there (codeml.c:357), and one engine per gene here.

Two ways to run it.  ``scan``: genes in parallel on a pool of host threads,
each owning one engine, so one HIP stream, at a time: the optimiser loop
runs in the library without the interpreter lock, and the small per-gene
kernels of different genes overlap on the GPU; work is handed out
largest-first (npatt x taxa), the LPT order, so the long genes do not end
the scan alone.  ``scan_batched``: genes of similar pattern count grouped
into batches (one forest of genes per batch, bsa_batch_*), each batch's H1
fits and then its H0 fits served in rounds of the genes whose points are
ready, so one launch per tree level serves many genes; a few batches run
at once on host threads.
Small genes are bound by launch processing, not by the GPU's throughput,
which the batches amortise.

The synthetic genome follows the 2,000-gene scan workload: Yule trees of
12-64 taxa with one random foreground branch, codon counts lognormal with
median 400 and sigma 0.7, clipped to [50, 5000], parameters drawn as
kappa ~ U(1, 4), w0 ~ U(0.01, 0.5), w2 ~ U(1.5, 6), p0 = 0.5, p1 = 0.3.
"""
import time
from concurrent.futures import ThreadPoolExecutor, as_completed
from dataclasses import dataclass

import numpy as np

from . import codeml, engine, formats, simulate

SCAN_SIGMA = 0.7
SCAN_MEDIAN = 400
SCAN_CLIP = (50, 5000)
SCAN_TAXA = (12, 64)


@dataclass
class Gene:
    name: str
    tree: formats.Tree
    data: formats.CodonData
    truth: dict


def simulate_genes(n, seed=2024, sigma=SCAN_SIGMA, median=SCAN_MEDIAN, clip=SCAN_CLIP,
                   taxa=SCAN_TAXA, mean_bl=0.05):
    """n synthetic genes of the genome-scan workload (seeded)."""
    rng = np.random.default_rng(seed)
    genes = []
    for g in range(n):
        ns = int(rng.integers(taxa[0], taxa[1] + 1))
        nc = int(np.clip(round(median * np.exp(sigma * rng.normal())), clip[0], clip[1]))
        p = simulate.uniform_params(rng)
        tree = simulate.yule_tree(ns, mean_bl, rng)
        _, pi = simulate.dirichlet_f3x4(rng)
        seqs, _ = simulate.simulate_model_a(tree, nc, p["kappa"], p["w0"], p["w2"], p["p0"],
                                            p["p1"], pi, rng)
        data = formats.load_codon_data(tree.names, seqs)
        genes.append(Gene(f"gene{g + 1:05d}", tree, data, dict(p, ns=ns, ncodon=nc)))
    return genes


def fit_gene(gene, precision="fp64", gradient="adjoint", device=0, e=1e-8, graph=False):
    """H1 and H0 fits of one gene on its own engine, and the LRT.  With
    ``graph`` each evaluation's device work is replayed from a HIP graph
    (measured slower on MI355X once several genes run at once, so off)."""
    t0 = time.perf_counter()
    # a tree file without branch lengths: every branch starts at 0.1
    start = codeml.default_start(gene.tree, 2.0, 1.5, False)
    start[:gene.tree.nbranch] = 0.1
    eng = engine.BranchSiteEngine(gene.tree, gene.data, precision=precision, device=device)
    eng.use_graph(graph)
    try:
        h1 = codeml.fit_model_a(gene.tree, gene.data, False, 1.5, start=start,
                                gradient=gradient, e=e, eng=eng)
        h0 = codeml.fit_model_a(gene.tree, gene.data, True, 1.0, start=h1.x[:-1],
                                gradient=gradient, e=e, eng=eng)
        if h0.lnL > h1.lnL + 1e-6:
            # H1 contains H0 (w2 = 1): a lower H1 maximum is a local one; fit
            # H1 again from the H0 estimates
            h1b = codeml.fit_model_a(gene.tree, gene.data, False, 1.5,
                                     start=np.append(h0.x, 1.0), gradient=gradient, e=e,
                                     eng=eng)
            h1b.iterations += h1.iterations
            h1b.evaluations += h1.evaluations
            h1 = h1b if h1b.lnL > h1.lnL else h1
    finally:
        eng.close()
    r = codeml.lrt(h0, h1)
    return {"gene": gene.name, "ns": gene.tree.ns, "npatt": gene.data.npatt,
            "lnL1": h1.lnL, "lnL0": h0.lnL, "stat": r["stat"], "p": r["p_mixture"],
            "w2": h1.w2, "x1": h1.x.tolist(),
            "params1": [h1.kappa, h1.p0, h1.p1, h1.w0, h1.w2],
            "iterations": h1.iterations + h0.iterations,
            "evaluations": h1.evaluations + h0.evaluations,
            "x0": h0.x.tolist(), "converged": h1.converged and h0.converged,
            "stop": (h1.stop, h0.stop), "seconds": time.perf_counter() - t0}


def scan(genes, workers=8, precision="fp64", gradient="adjoint", device=0, e=1e-8, graph=False):
    """Fit every gene (H1, H0, LRT) with `workers` concurrent engines.
    Returns (results in input order, wall seconds)."""
    order = sorted(range(len(genes)),
                   key=lambda i: -genes[i].data.npatt * genes[i].tree.ns)
    out = [None] * len(genes)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        futs = {i: ex.submit(fit_gene, genes[i], precision, gradient, device, e, graph)
                for i in order}
        for i, f in futs.items():
            out[i] = f.result()
    return out, time.perf_counter() - t0


def _pad(n, p=128):
    return (n + p - 1) // p * p


def batch_groups(genes, batch_genes=64, max_waste=0.25):
    """Gene indices in batches: largest pattern count first; a batch takes
    genes while their padded pattern count is within `max_waste` of its
    first gene's (the batch pads every gene to that count)."""
    order = sorted(range(len(genes)), key=lambda i: -genes[i].data.npatt)
    groups, cur = [], []
    for i in order:
        if cur and (len(cur) >= batch_genes or
                    _pad(genes[i].data.npatt) < (1 - max_waste) * _pad(genes[cur[0]].data.npatt)):
            groups.append(cur)
            cur = []
        cur.append(i)
    if cur:
        groups.append(cur)
    return groups


def _finish_fp64(genes, x1, x0, device, e):
    """FP64 finish of FP32 batch fits: H1 from the FP32 H1 estimates and H0
    from the FP32 H0 estimates, on an FP64 batch engine
    (codeml.finish_fp64_batch: Newton iterations on the adjoint gradient's
    differences, BFGS where they do not converge), then H1 again from the H0
    estimates where H0 comes out above H1 (as bsa_batch_fit_lrt does).
    FP32 partials carry rounding noise of ~1e-9 |lnL| that ends a fit at
    that floor; the finish takes the estimates to the FP64 maximum.
    Returns (x1, x0, -lnL (n, 2), H1 params, per-gene finish info)."""
    n = len(genes)
    b = engine.BranchSiteBatch([(g.tree, g.data) for g in genes], precision="fp64",
                               device=device)
    try:
        b.set_model(-1, True, 1.0)
        x0f, f0, i0 = codeml.finish_fp64_batch(b, x0, [None] * n, [None] * n, e)
        b.set_model(-1, False)
        x1f, f1, i1 = codeml.finish_fp64_batch(b, x1, [None] * n, [None] * n, e)
        refit = [i for i in range(n) if -f0[i] > -f1[i] + 1e-6]
        if refit:
            # one batch fit; the other genes are held at their maximum (lo = hi)
            xs = [np.append(x0f[i], 1.0) if i in refit else x1f[i] for i in range(n)]
            held = [None if i in refit else x1f[i] for i in range(n)]
            xb, fb, ib = b.fit(xs, e=e, lo=held, hi=held)
            for i in refit:
                i1[i] = codeml._merge_info(i1[i], ib[i])
                if fb[i] < f1[i]:
                    x1f[i], f1[i] = xb[i], fb[i]
        b.lnl(x1f)   # every gene's last evaluation at its H1 estimates
        par1 = [b.last_params(i) for i in range(n)]
    finally:
        b.close()
    info = [{"h1": i1[i], "h0": i0[i], "refit": i in refit} for i in range(n)]
    return x1f, x0f, np.stack([f1, f0], axis=1), par1, info


def batch_timing(b, ngenes):
    """Device-side record of a batch engine's fits from its own event timers
    (bsa_batch_timing): rounds served, device time per phase summed over
    the rounds, wall time of the fits, launches per round."""
    tm = b.timing()
    rounds = max(tm["n_eval"], 1.0)
    phases = ("pmat", "prune", "root", "adjoint", "branch", "finish")
    dev = sum(tm["sum_" + k] for k in phases)
    # per round: P(t) + tip tables (2), pruning levels, root + sums (2), child
    # adjoints per depth + branch matrices, adjoint finish (3)
    launches = 2 + tm["n_launch"] / rounds + 2 + tm["n_adj_launch"] / rounds + 3
    return {"genes": ngenes, "rounds": int(tm["n_eval"]), "wall_ms": tm["sum_wall"],
            "device_ms": dev, "device_busy": dev / tm["sum_wall"] if tm["sum_wall"] > 0 else 0.0,
            "device_ms_per_round": dev / rounds, "wall_ms_per_round": tm["sum_wall"] / rounds,
            "launches_per_round": launches,
            "phase_ms": {k: tm["sum_" + k] for k in phases}}


def fit_batch(genes, precision="fp64", device=0, e=1e-8, graph=False, stats=None,
              prune="levels"):
    """H1 and H0 fits and the LRT of a list of genes on one batch engine
    (bsa_batch_fit_lrt): the sequence of ``fit_gene`` (H1, H0 from the H1
    estimates, H1 again from the H0 estimates where H0 lies above H1), each
    gene's chain in its own optimiser thread, so no gene waits for another
    gene's fit.  With precision "fp32" the chains run on FP32 partials and
    are finished on FP64 ones (``_finish_fp64``); lnL, estimates and the
    statistic are then the FP64 ones, as in ``fit_gene``.  Returns the
    per-gene dicts of ``fit_gene`` (seconds = the batch's wall time).  With
    a list `stats`, the batch's ``batch_timing`` record of its chained fits
    is appended to it.  `prune`: the batch engine's pruning mode ("levels"
    or "walk", bsa_batch_set_prune)."""
    t0 = time.perf_counter()
    n = len(genes)
    b = engine.BranchSiteBatch([(g.tree, g.data) for g in genes], precision=precision,
                               device=device)
    b.use_graph(graph)
    try:
        b.set_prune(prune)
        starts = []
        for g in genes:
            s = codeml.default_start(g.tree, 2.0, 1.5, False)
            s[:g.tree.nbranch] = 0.1
            starts.append(s)
        b.timing(reset=True)
        x1, x0, negl, info = b.fit_lrt(starts, e=e)
        if stats is not None:
            st = batch_timing(b, n)
            st["evaluations"] = int(sum(inf["h1_evaluations"] + inf["h0_evaluations"]
                                        for inf in info))
            st["genes_per_round"] = st["evaluations"] / max(st["rounds"], 1)
            stats.append(st)
        par1 = [b.last_params(i) for i in range(n)]
    finally:
        b.close()
    stages = [{"h1": {k[3:]: v for k, v in inf.items() if k.startswith("h1_")},
               "h0": {k[3:]: v for k, v in inf.items() if k.startswith("h0_")},
               "refit": bool(inf["refit"])} for inf in info]
    first = stages
    if precision == "fp32":
        x1, x0, negl, par1, fin = _finish_fp64(genes, x1, x0, device, e)
        stages = [{"h1": codeml._merge_info(a["h1"], c["h1"]),
                   "h0": codeml._merge_info(a["h0"], c["h0"]),
                   "refit": a["refit"] or c["refit"]} for a, c in zip(stages, fin)]
    wall = time.perf_counter() - t0
    out = []
    for i, g in enumerate(genes):
        st = stages[i]
        h1 = codeml.Fit("H1", -negl[i, 0], x1[i], *par1[i][:5], par1[i][7:11].copy(),
                        x1[i][:g.tree.nbranch].copy(), len(x1[i]), st["h1"]["iterations"],
                        st["h1"]["evaluations"], st["h1"]["converged"], wall, {},
                        st["h1"]["stop"])
        h0 = codeml.Fit("H0", -negl[i, 1], x0[i], 0, 0, 0, 0, 1.0, None, None, len(x0[i]),
                        st["h0"]["iterations"], st["h0"]["evaluations"], st["h0"]["converged"],
                        wall, {}, st["h0"]["stop"])
        r = codeml.lrt(h0, h1)
        d = {"gene": g.name, "ns": g.tree.ns, "npatt": g.data.npatt,
             "lnL1": h1.lnL, "lnL0": h0.lnL, "stat": r["stat"], "p": r["p_mixture"],
             "w2": h1.w2, "x1": h1.x.tolist(), "x0": h0.x.tolist(),
             "params1": [h1.kappa, h1.p0, h1.p1, h1.w0, h1.w2],
             "iterations": h1.iterations + h0.iterations,
             "evaluations": h1.evaluations + h0.evaluations,
             "converged": h1.converged and h0.converged, "stop": (h1.stop, h0.stop),
             "refit": st["refit"], "precision": precision, "seconds": wall}
        if precision == "fp32":
            d["fp32_evaluations"] = first[i]["h1"]["evaluations"] + first[i]["h0"]["evaluations"]
        out.append(d)
    return out


def scan_batched(genes, batch_genes=64, concurrent=6, precision="fp64", device=0, e=1e-8,
                 graph=False, max_waste=0.25, stats=None, progress=None, prune="levels"):
    """Fit every gene (H1, H0, LRT) in batches of up to `batch_genes` genes
    of similar size, `concurrent` batches at a time, each batch engine
    pruning by `prune` ("levels" or "walk").  Returns (results in input
    order, wall seconds); `stats` (a list) receives each batch's
    ``batch_timing`` record, and progress(batches done, batches, genes of
    the batch) is called as each batch ends."""
    groups = batch_groups(genes, batch_genes, max_waste)
    out = [None] * len(genes)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, concurrent)) as ex:
        futs = {ex.submit(fit_batch, [genes[i] for i in grp], precision, device, e, graph,
                          stats, prune): grp for grp in groups}
        for k, f in enumerate(as_completed(futs), 1):
            grp = futs[f]
            for i, r in zip(grp, f.result()):
                out[i] = r
            if progress:
                progress(k, len(groups), len(grp))
    return out, time.perf_counter() - t0
