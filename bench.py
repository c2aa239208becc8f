#!/usr/bin/env python3
"""Throughput of the branch-site model A likelihood path on one MI355X.

Metric: lnL+gradient evaluations per second at 128 taxa x 10,000 codons.
One step = one evaluation of -lnL and its gradient in codeml's parameter
vector (np = 2*128-3 + 5 = 258 parameters).  The timed path is the one the
engine ships for optimisation: FP32-stored conditional likelihoods (FP64
arithmetic, stochastic rounding) with the adjoint gradient
(bsa_lnl_grad: forward pruning + one preorder adjoint sweep).  Beside it, in
the same run, the plain FP64 path with ming2's forward-difference gradient
(gradientB: np+1 = 259 likelihood evaluations per step) is timed on the same
GPU, and the reference's own CPU codeml (built from its sources into
oracle/_ref/codeml) on the same alignment at fixed parameters, its set-up
timed apart and subtracted.

Inputs are synthetic (seeded simulation under model A, parameter draws of
the 128-taxon gradient-check configuration) and resident in HBM before
timing starts.
"""
import argparse
import collections
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cucodeml_amd import engine, formats, simulate  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def make_workload(ns, ncodon, seed, branch_mean=0.05, gap=0.0):
    """Seeded Yule tree (branch lengths exponential of the given mean) and a
    model-A alignment simulated on it, with a fraction `gap` of the codons
    missing (512 taxa, 50,000 codons, mean 0.02, gap 0.1, seed 4 is the
    large-gene workload, the same draws as tests/test_gpu_large.py)."""
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, branch_mean, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=gap)
    data = formats.load_codon_data(tree.names, seqs)
    return tree, data, p, seqs


CTL = """seqfile = seq.phy
treefile = tree.nwk
outfile = out
noisy = 0
verbose = 0
runmode = 0
seqtype = 1
CodonFreq = 2
clock = 0
model = 2
NSsites = 2
icode = 0
fix_kappa = 0
kappa = 2
fix_omega = 1
omega = {w2}
fix_alpha = 1
alpha = 0
getSE = 0
RateAncestor = 0
cleandata = 0
method = 0
"""


REF_CODEML = os.path.join(ROOT, "oracle", "_ref", "codeml")


def ref_run(names, seqs, newick, ctl, xin=None, timeout=3600, log=None):
    """One run of the reference CPU codeml in a scratch folder: (wall
    seconds, stdout, `out` text), or None if it failed.  With `log` (a
    path) its standard output goes to that file as it runs."""
    work = tempfile.mkdtemp(prefix="bsa_cpu_")
    try:
        formats.write_phylip(os.path.join(work, "seq.phy"), names, seqs)
        with open(os.path.join(work, "tree.nwk"), "w") as f:
            f.write(f"{len(names)} 1\n{newick}\n")
        if xin is not None:
            with open(os.path.join(work, "in.codeml"), "w") as f:
                f.write(xin + "\n")
        with open(os.path.join(work, "codeml.ctl"), "w") as f:
            f.write(ctl)
        t0 = time.perf_counter()
        if log:
            with open(log, "w") as lf:
                r = subprocess.run([REF_CODEML, "codeml.ctl"], cwd=work, stdout=lf,
                                   stderr=subprocess.STDOUT, text=True, input="\n",
                                   timeout=timeout)
            stdout = ""
        else:
            r = subprocess.run([REF_CODEML, "codeml.ctl"], cwd=work, capture_output=True,
                               text=True, input="\n", timeout=timeout)
            stdout = r.stdout
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        with open(os.path.join(work, "out")) as f:
            out = f.read()
        return wall, stdout, out
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cpu_info():
    """CPU model of this host (for the baseline records)"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tree, seqs, p, npatt_full, np_full, sample_codons):
    """Reference CPU codeml (one thread) at fixed parameters (in.codeml -1).
    One run makes 3 full likelihood evaluations (codeml.c:670, 5107, 833);
    a second run on the first 10 codons of the same alignment gives the
    set-up (input, distances, eigen systems, and its own three evaluations
    of 10 codons), which is subtracted.  The sample is the first
    `sample_codons` codons (the whole alignment by default), its per-lnL
    time scaled by patterns to the full alignment when it is smaller."""
    if not os.path.exists(REF_CODEML):
        return None
    x = list(tree.branch_vector()) + [p["kappa"], p["p0"], p["p1"], p["w0"]]
    xin = "-1 " + " ".join(f"{v:.12f}" for v in x)
    ctl = CTL.format(w2=p["w2"])
    nwk = tree.newick(digits=10)
    sub = [s[:3 * sample_codons] for s in seqs]
    d_sub = formats.load_codon_data(tree.names, sub)
    run = ref_run(tree.names, sub, nwk, ctl, xin)
    setup = ref_run(tree.names, [s[:30] for s in seqs], nwk, ctl, xin)
    if run is None or setup is None:
        return None
    wall, wall0 = run[0], setup[0]
    per_lnl_sample = (wall - wall0) / 3.0
    per_lnl_full = per_lnl_sample * npatt_full / d_sub.npatt
    evals_per_s = 1.0 / ((np_full + 1) * per_lnl_full)
    return {"value": evals_per_s, "unit": "lnL+gradient evals/s", "cores": 1,
            "kind": "reference", "cpu": cpu_info(),
            "sample": (f"reference codeml (CPU, gcc -O3, 1 thread) at fixed parameters on "
                       f"{tree.ns} taxa x {len(sub[0]) // 3} codons ({d_sub.npatt} patterns): "
                       f"{wall:.2f} s per run of 3 lnL evaluations, {wall0:.2f} s for the "
                       f"set-up run (10 codons) -> {per_lnl_sample:.3f} s/lnL"
                       + (f", scaled by patterns to {npatt_full} ({per_lnl_full:.3f} s/lnL)"
                          if d_sub.npatt != npatt_full else "")
                       + f"; x (np+1) = {np_full + 1} evaluations per forward-difference "
                         f"gradient (ming2's gradientB)")}


def cpu_scan_baseline(ngenes, nproc, logdir=None):
    """genes/hour of the reference CPU codeml on the genome-scan workload, as
    a scan runs it: one single-thread process per gene and hypothesis (H1,
    then H0), `nproc` genes at once on `nproc` cores, each from codeml's own
    starting values (tree without branch lengths).  The sample is the nproc
    genes of the first `ngenes` whose work (patterns x taxa) is nearest the
    median: the workload's sizes are lognormal, so the mean gene costs more
    and the figure favours the CPU.  With `logdir` each run writes its
    iteration log there (noisy = 3) as it goes."""
    if not os.path.exists(REF_CODEML):
        return None
    from concurrent.futures import ThreadPoolExecutor
    from cucodeml_amd import scan
    genes = scan.simulate_genes(ngenes)
    work = np.array([g.data.npatt * g.tree.ns for g in genes], dtype=float)
    med = float(np.median(work))
    pick = sorted(np.argsort(np.abs(work - med))[:nproc].tolist())
    ctl1 = CTL.replace("fix_omega = 1", "fix_omega = 0").format(w2=1.5)
    ctl0 = CTL.format(w2=1.0)
    if logdir:
        os.makedirs(logdir, exist_ok=True)
        ctl1, ctl0 = (c.replace("noisy = 0", "noisy = 3") for c in (ctl1, ctl0))

    def fit(i):
        g = genes[i]
        seqs = formats.sequences_of(g.data)
        nwk = g.tree.newick(lengths=False)
        t0 = time.perf_counter()
        logs = [os.path.join(logdir, f"{g.name}_{h}.log") if logdir else None
                for h in ("H1", "H0")]
        h1 = ref_run(g.tree.names, seqs, nwk, ctl1, timeout=6 * 3600, log=logs[0])
        h0 = ref_run(g.tree.names, seqs, nwk, ctl0, timeout=6 * 3600, log=logs[1])
        el = time.perf_counter() - t0
        print(f"reference fits of {g.name} ({g.tree.ns} taxa x {g.data.npatt} patterns): "
              f"{el:.0f} s", file=sys.stderr, flush=True)
        ok = h1 is not None and h0 is not None
        return {"gene": g.name, "ns": g.tree.ns, "npatt": g.data.npatt, "seconds": el, "ok": ok}

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=nproc) as ex:
        res = list(ex.map(fit, pick))
    wall = time.perf_counter() - t0
    return {"value": len(pick) / wall * 3600.0, "unit": "genes/hour", "cores": nproc,
            "kind": "reference", "cpu": cpu_info(), "wall_s": wall,
            "all_ok": all(r["ok"] for r in res), "genes": res,
            "sample": (f"{len(pick)} genes of the first {ngenes} of the scan workload nearest "
                       f"its median work ({med:.0f} patterns x taxa), one reference codeml "
                       f"process per gene and hypothesis (H1, then H0) from codeml's own "
                       f"starting values, {nproc} at once")}


def roofline(work_bytes, work_flops, ms_sum, nev, launches, kernel, matrix="fp64"):
    """Roofline entry of one kernel from its algorithmic work per evaluation
    and its summed HIP-event time: bound = the larger of the HBM time
    (bytes / 8 TB/s) and the matrix-core time (flops / the dense peak of the
    kernel's matrix-core arithmetic: FP64 78.6 or FP32 157.3 TFLOP/s)."""
    peak_tf = FP32_PEAK_TFLOPS if matrix == "fp32" else FP64_PEAK_TFLOPS
    sec = ms_sum / 1e3 / nev
    gbs = work_bytes / sec / 1e9
    tf = work_flops / sec / 1e12
    mfma = work_flops / (peak_tf * 1e12) > work_bytes / (HBM_PEAK_GBS * 1e9)
    ach, peak, unit = (tf, peak_tf, "TFLOP/s") if mfma else (gbs, HBM_PEAK_GBS, "GB/s")
    return {"bound": "mfma" if mfma else "hbm", "achieved": ach, "peak": peak, "unit": unit,
            "frac": ach / peak, "traffic": None, "kernel": kernel,
            "bytes_per_eval": work_bytes, "flops_per_eval": work_flops,
            "avg_launch_us": ms_sum * 1e3 / max(launches, 1), "launches_per_eval": launches / nev,
            "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "matrix": matrix,
            "matrix_tflops": tf, "matrix_frac": tf / peak_tf}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` measured by the rocprofv3 PMC runs
    summarised under profiles/ for this same workload, or None."""
    import glob
    short = re.split(r"[<\s]", kernel)[0]
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+([0-9.eE+]+)", line)
                if m and m.group(1) == short:
                    return float(m.group(2)), os.path.relpath(path, ROOT)
    return None


def genome_scan(ngenes, workers, precision, batch_genes, concurrent):
    """genes/hour of the genome scan (H1 + H0 fits and LRT per gene) on a
    sample of the 2,000-gene workload: genes of similar size batched into
    forests (batch_genes > 0; `concurrent` forests at a time) or one engine
    per gene on `workers` threads (batch_genes = 0)."""
    from cucodeml_amd import scan
    genes = scan.simulate_genes(ngenes)
    if batch_genes > 0:
        scan.scan_batched(genes[:min(4, ngenes)], batch_genes=4, concurrent=1,
                          precision=precision)   # warm-up
        stats = []
        t0 = time.perf_counter()

        def progress(k, n, m):
            print(f"scan: batch {k}/{n} ({m} genes) done at {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)

        res, wall = scan.scan_batched(genes, batch_genes=batch_genes, concurrent=concurrent,
                                      precision=precision, stats=stats, progress=progress)
        how = {"mode": "batched", "batch_genes": batch_genes, "concurrent_batches": concurrent,
               "device": scan_device_record(stats, wall)}
    else:
        scan.scan(genes[:min(2, ngenes)], workers=1, precision=precision)   # warm-up
        res, wall = scan.scan(genes, workers=workers, precision=precision)
        how = {"mode": "engine per gene", "workers": workers}
    return {"value": ngenes / wall * 3600.0, "unit": "genes/hour",
            "sample": (f"first {ngenes} genes of the seeded 2,000-gene scan workload (Yule "
                       f"trees of {scan.SCAN_TAXA[0]}-{scan.SCAN_TAXA[1]} taxa, lognormal codon "
                       f"counts median {scan.SCAN_MEDIAN}, sigma {scan.SCAN_SIGMA}, clipped to "
                       f"{list(scan.SCAN_CLIP)})"),
            "per_gene": "H1 fit + H0 fit (bounded BFGS, adjoint gradient) + LRT",
            "precision": precision, **how, "wall_s": wall,
            "converged": int(sum(r["converged"] for r in res)),
            "stops": dict(collections.Counter(f"{r['stop'][0]}/{r['stop'][1]}" for r in res)),
            "mean_evaluations": float(np.mean([r["evaluations"] for r in res]))}


def batch_pruning(ngenes, taxa=24, codons=300, reps=3):
    """Whole-batch lnL + gradient evaluations of `ngenes` synthetic genes
    (seeded Yule trees of `taxa` taxa, `codons` codons, FP64 partials) with
    the pruning by level launches and by the persistent walk: device ms per
    evaluation (events), and whether the two give the same results bitwise."""
    rng = np.random.default_rng(5)
    genes, xs = [], []
    for _ in range(ngenes):
        tree = simulate.yule_tree(taxa, 0.05, rng)
        _, pi = simulate.dirichlet_f3x4(rng)
        seqs, _ = simulate.simulate_model_a(tree, codons, 2.0, 0.2, 3.0, 0.5, 0.3, pi, rng)
        genes.append((tree, formats.load_codon_data(tree.names, seqs)))
        xs.append(engine.x_from_params(tree.branch_vector(), 2.0, 0.5, 0.3, 0.2, 3.0))
    out = {"workload": f"{ngenes} genes x {taxa} taxa x {codons} codons, fp64 partials, "
                       f"lnL + adjoint gradient of every gene per evaluation"}
    res = {}
    phases = ("pmat", "prune", "root", "adjoint", "branch", "finish")
    for mode in ("levels", "walk"):
        b = engine.BranchSiteBatch(genes, precision="fp64")
        try:
            b.set_prune(mode)
            b.lnl_grad(xs)
            b.timing(reset=True)
            for _ in range(reps):
                f, g = b.lnl_grad(xs)
            tm = b.timing()
        finally:
            b.close()
        res[mode] = (f, g)
        out[mode] = {"prune_ms": tm["sum_prune"] / reps,
                     "device_ms": sum(tm["sum_" + k] for k in phases) / reps,
                     "pruning_launches": tm["n_launch"] / reps}
    out["bitwise_equal"] = bool(np.array_equal(res["levels"][0], res["walk"][0]) and
                                all(np.array_equal(a, c) for a, c in
                                    zip(res["levels"][1], res["walk"][1])))
    return out


def scan_device_record(stats, wall):
    """The batch engines' own event timers summed over a scan's batches:
    each batch engine serves its genes' evaluation requests in rounds on its
    own stream (batches run at once on separate streams)."""
    rounds = sum(st["rounds"] for st in stats)
    dev = sum(st["device_ms"] for st in stats)
    bwall = sum(st["wall_ms"] for st in stats)
    return {"batches": len(stats), "rounds": rounds,
            "device_ms_sum": dev,
            "device_busy_per_batch": dev / bwall if bwall > 0 else 0.0,
            "stream_time_over_scan_wall": dev / (wall * 1e3) if wall > 0 else 0.0,
            "device_ms_per_round": dev / max(rounds, 1),
            "wall_ms_per_round": bwall / max(rounds, 1),
            "host_ms_per_round": (bwall - dev) / max(rounds, 1),
            "launches_per_round": sum(st["launches_per_round"] * st["rounds"] for st in stats)
            / max(rounds, 1),
            "genes_per_round": sum(st["evaluations"] for st in stats) / max(rounds, 1),
            "phase_ms": {k: sum(st["phase_ms"][k] for st in stats)
                         for k in stats[0]["phase_ms"]} if stats else {}}


def large_fit(ns=512, ncodon=50000, gap=0.1, seed=4, e=1e-8):
    """Full ML fit (H1) of the large-gene workload, 512 taxa x 50,000 codons
    with 10 % gaps, from codeml's start (all branches 0.1), by the driver's
    own fit (codeml.fit_model_a): with FP64 partials; with FP32 partials
    finished on FP64; and, from the FP64 estimates, the FP32 partials' own
    fit, which shows where FP32 partials put the same maximum."""
    from cucodeml_amd import codeml
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, 0.02, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=gap)
    data = formats.load_codon_data(tree.names, seqs)
    x0 = codeml.default_start(tree, 2.0, 1.5, False)
    x0[:tree.nbranch] = 0.1
    names = ["kappa", "p0", "p1", "w0", "w2"]

    def fit(precision, x, label, polish=True):
        t0 = time.perf_counter()
        r = codeml.fit_model_a(tree, data, precision=precision, start=x, e=e, polish=polish)
        dt = time.perf_counter() - t0
        print(f"large fit {label}: {r.evaluations} evaluations, {r.iterations} iterations, "
              f"-lnL {-r.lnL:.6f}, {dt:.1f} s", file=sys.stderr, flush=True)
        return r, dt

    def est(r):
        return dict(zip(names, [r.kappa, r.p0, r.p1, r.w0, r.w2]))

    r64, t64 = fit("fp64", x0, "fp64")
    r32, t32 = fit("fp32", x0, "fp32 finished on fp64")
    # the FP32 fit started at the FP64 estimates: where FP32 partials put
    # the same maximum
    r32m, _ = fit("fp32", r64.x, "fp32 from the fp64 estimates", polish=False)
    e64 = engine.BranchSiteEngine(tree, data, precision="fp64")
    try:
        f32m_64 = e64.lnl(r32m.x)
    finally:
        e64.close()
    nb = tree.nbranch
    par = lambda r: np.array([r.kappa, r.p0, r.p1, r.w0, r.w2])   # noqa: E731
    ex = r32.extra
    return {"workload": f"{ns} taxa x {ncodon} codons ({data.npatt} patterns), {gap:.0%} gaps, "
                        f"H1 from all branches 0.1", "e": e,
            "fp64": {"neg_lnl": -r64.lnL, "evaluations": r64.evaluations,
                     "iterations": r64.iterations, "stop": r64.stop, "seconds": t64,
                     "estimates": est(r64)},
            "fp32_then_fp64": {"neg_lnl_fp32": -ex["fp32_lnL"], "fp32_stop": ex["fp32_stop"],
                               "fp32_evaluations": ex["fp32_evaluations"],
                               "neg_lnl_after_fp64_finish": -r32.lnL,
                               "evaluations": r32.evaluations, "iterations": r32.iterations,
                               "stop": r32.stop, "seconds": t32,
                               "estimates": est(r32),
                               "delta_neg_lnl_vs_fp64": r64.lnL - r32.lnL},
            "fp32_at_the_fp64_maximum": {
                "neg_lnl_fp32": -r32m.lnL, "neg_lnl_fp64_at_its_estimates": f32m_64,
                "delta_neg_lnl_vs_fp64": f32m_64 + r64.lnL, "evaluations": r32m.evaluations,
                "max_abs_param_diff": float(np.abs(par(r32m) - par(r64)).max()),
                "max_abs_branch_diff": float(np.abs(r32m.x[:nb] - r64.x[:nb]).max())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--codons", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--branch-mean", type=float, default=0.05,
                    help="mean branch length of the simulated tree")
    ap.add_argument("--gap", type=float, default=0.0,
                    help="fraction of missing codons in the simulated alignment")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--gradient", choices=["adjoint", "fd"], default="adjoint")
    ap.add_argument("--kernel", choices=["mfma", "mfma32", "fp32mc", "valu"], default=None,
                    help="default: mfma32 with fp32 partials (FP32 matrix-core child adjoints), "
                         "mfma with fp64; fp32mc: the pruning on the FP32 matrix cores too")
    ap.add_argument("--tile", type=int, default=0, help="patterns per pruning block (64/128)")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--large-fit", action="store_true",
                    help="also run the full H1 fit of 512 taxa x 50,000 codons, FP32 and FP64 "
                         "(minutes; progress on stderr)")
    ap.add_argument("--large-fit-e", type=float, default=1e-8,
                    help="relative convergence tolerance of the large fit")
    ap.add_argument("--fd-steps", type=int, default=1,
                    help="steps of the FP64 forward-difference path timed beside (0: skip)")
    ap.add_argument("--cpu-sample-codons", type=int, default=0,
                    help="codons of the reference CPU baseline's sample (0: all)")
    ap.add_argument("--scan-precision", choices=["fp64", "fp32"], default="fp64",
                    help="partials of the scan's fits (fp32: FP32 chains finished on FP64)")
    ap.add_argument("--scan-cpu", type=int, default=0,
                    help="also time the reference CPU codeml on this many scan genes at once "
                         "(one core each; full H1 + H0 fits: tens of minutes)")
    ap.add_argument("--scan-cpu-log", default=None,
                    help="folder for the reference runs' iteration logs (--scan-cpu)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scan-genes", type=int, default=2000,
                    help="genes of the genome scan timed beside: the first N of the seeded "
                         "2,000-gene workload (default: all of it, about a minute; 0: skip)")
    ap.add_argument("--scan-batch", type=int, default=256,
                    help="genes per batch engine in the scan (0: one engine per gene)")
    ap.add_argument("--scan-concurrent", type=int, default=6,
                    help="batches fitted at once in the scan")
    ap.add_argument("--scan-workers", type=int, default=8,
                    help="engines at once when the scan runs one engine per gene")
    ap.add_argument("--batch-pruning", type=int, default=0,
                    help="also time whole-batch evaluations of this many small genes with "
                         "level launches and with the persistent pruning (0: skip)")
    ap.add_argument("--no-beb", dest="beb", action="store_false",
                    help="skip timing one Bayes empirical Bayes analysis beside")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU; sharding a gene's site patterns "
              "over several GPUs is not built", file=sys.stderr)
        sys.exit(2)

    args.kernel = args.kernel or engine.default_kernel(args.precision)
    tree, data, p, seqs = make_workload(args.taxa, args.codons, args.seed, args.branch_mean,
                                        args.gap)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                             p["w2"])
    npar = len(x)
    eng = engine.BranchSiteEngine(tree, data, precision=args.precision, kernel=args.kernel,
                                  tile=args.tile)
    step = eng.lnl_grad if args.gradient == "adjoint" else eng.gradientB
    # every step is at a new parameter vector, as in an optimisation: the
    # rate-matrix eigen systems (host) and all P(t) are recomputed each time
    xs = [x * (1.0 + 1e-7 * (k + 1)) for k in range(args.warmup + args.steps)]
    for k in range(args.warmup):
        step(xs[k])
    # launch shapes picked by the engine's autotuner (untimed, like warm-up)
    shapes = eng.tune(xs[max(0, args.warmup - 1)]) \
        if args.tune and args.gradient == "adjoint" else None
    eng.timing(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        f, g = step(xs[args.warmup + k])
    elapsed = time.perf_counter() - t0
    x = xs[-1]
    tm = eng.timing(reset=True)
    work = eng.work()
    nev = tm["n_eval"]
    sto = "double" if args.precision == "fp64" else "float"
    if args.kernel == "fp32mc":
        prune_rl = roofline(work[1], work[0], tm["sum_prune"], nev, tm["n_launch"],
                            "k_prune32 (pruning levels on the FP32 matrix cores, each row's two "
                            "largest child terms and the identity term in FP64)", "fp32")
    else:
        prune_rl = roofline(work[1], work[0], tm["sum_prune"], nev, tm["n_launch"],
                            (f"k_prune_res/k_prune_mfma<{sto}> (pruning levels, FP64 arithmetic, "
                             f"{sto} storage)") if args.kernel != "valu" else
                            f"k_prune_valu<{sto}> (FP64 arithmetic, {sto} storage)")
    rls = [(tm["sum_prune"], prune_rl)]
    adj_rl = br_rl = None
    if args.gradient == "adjoint":
        nlev = tm["n_adj_launch"] - nev   # child-adjoint launches (one branch-matrix launch each)
        m32 = args.kernel in ("mfma32", "fp32mc")
        adj_rl = roofline(work[7], work[6], tm["sum_adjoint"], nev, nlev,
                          "k_adj_a32 (child adjoints, FP32 matrix-core products, identity term "
                          "in FP64)" if m32 else
                          f"k_adj_a<{sto}> (child adjoints, FP64 arithmetic)",
                          "fp32" if m32 else "fp64")
        # the branch matrices' flops are the interior children's products
        br_rl = roofline(work[9], work[8], tm["sum_branch"], nev, nev,
                         "k_adj_m<float, true> (branch matrices: interior children on the FP32 "
                         "matrix cores in 16-pattern groups summed in FP64, tips FP64)"
                         if m32 else
                         f"k_adj_m<{sto}> (branch matrices, FP64 arithmetic)",
                         "fp32" if m32 else "fp64")
        rls += [(tm["sum_adjoint"], adj_rl), (tm["sum_branch"], br_rl)]
    rl = max(rls, key=lambda r: r[0])[1]
    default_workload = (args.taxa, args.codons, args.seed, args.branch_mean, args.gap,
                        args.precision, args.gradient, args.kernel, args.tile,
                        args.tune) == (128, 10000, 2, 0.05, 0.0, "fp32", "adjoint", "mfma32", 0,
                                       True)
    tr = measured_traffic(rl["kernel"]) if default_workload else None
    if tr:
        rl["traffic"], rl["traffic_source"] = tr
    value = args.steps / elapsed
    out = {
        "metric": "lnL+gradient evals/sec at 128 taxa x 10k codons (branch-site model A)",
        "value": value,
        "unit": "lnL+gradient evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else (
            "f32-storage/f64-arith (child adjoints f32 matrix cores)" if args.kernel == "mfma32"
            else "f32 (pruning and gradient sweep on f32 matrix cores, f64 terms)"
            if args.kernel == "fp32mc" else "f32-storage/f64-arith"),
        "data": "synthetic: seeded model-A simulation (Yule tree, F3x4 Dirichlet(1))",
        "config": {"workload": f"branch-site model A, {args.taxa} taxa x {args.codons} codons "
                               f"({data.npatt} patterns), {args.precision} partials, "
                               f"{args.gradient} gradient of {npar} parameters",
                   "taxa": args.taxa, "codons": args.codons, "patterns": data.npatt,
                   "branch_mean": args.branch_mean, "gap": args.gap, "seed": args.seed,
                   "launch_shapes": shapes,
                   "np": npar, "gradient": args.gradient, "kernel": args.kernel,
                   "tile": args.tile or 64},
        "roofline": rl,
        "roofline_pruning": prune_rl,
        "roofline_child_adjoint": adj_rl,
        "roofline_branch_matrices": br_rl,
        "eval_ms": {k: tm["sum_" + k] / nev for k in
                    ("pmat", "prune", "root", "adjoint", "branch", "finish", "wall")},
        "lnL": -f,
    }
    if args.beb:
        eng.beb(x)
        t1 = time.perf_counter()
        eng.beb(x)
        out["beb"] = {"seconds": time.perf_counter() - t1, "precision": args.precision,
                      "what": "Bayes empirical Bayes site posteriors at x: 100 forward passes "
                              "(121 w sets) + 10^4-point grid integration"}
    if args.fd_steps > 0:
        e64 = engine.BranchSiteEngine(tree, data, precision="fp64",
                                      kernel="valu" if args.kernel == "valu" else "mfma")
        e64.gradientB(x)
        t1 = time.perf_counter()
        for _ in range(args.fd_steps):
            f64, _ = e64.gradientB(x)
        fd_el = (time.perf_counter() - t1) / args.fd_steps
        out["fd64_gpu"] = {"value": 1.0 / fd_el, "unit": "lnL+gradient evals/s",
                           "ms_per_step": fd_el * 1e3,
                           "path": f"FP64 partials, forward differences ({npar + 1} lnL/step)"}
        out["speedup_vs_fd64_gpu"] = value * fd_el
        out["lnL_error_vs_fp64"] = abs(f - f64)
    if args.scan_genes > 0:
        out["genome_scan"] = genome_scan(args.scan_genes, args.scan_workers, args.scan_precision,
                                         args.scan_batch, args.scan_concurrent)
    if args.batch_pruning > 0:
        out["batch_pruning"] = batch_pruning(args.batch_pruning)
    if args.large_fit:
        out["large_fit"] = large_fit(e=args.large_fit_e)
    if args.scan_cpu > 0:
        out["genome_scan_cpu_baseline"] = cpu_scan_baseline(max(args.scan_genes, 256),
                                                            args.scan_cpu, args.scan_cpu_log)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(tree, seqs, p, data.npatt, npar,
                                           args.cpu_sample_codons or args.codons)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
