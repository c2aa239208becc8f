"""bsa_lnl_noalloc, the entry point that allocates no device memory (every
buffer in the library's fixed per-GPU array), against bsa_lnl on engines of
the same genes: bitwise the same -lnL, per-class and per-site values, on two
calls in a row with the second on a smaller gene (the array keeps the first
call's contents, which must not leak into the second), in FP64 and FP32 and
under H0; and device free memory unchanged over calls after the first."""
import ctypes as C

import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

pytestmark = pytest.mark.gpu


def gene(seed, ns, ncodon, gaps=0.0):
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, 0.08, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=gaps)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    return tree, data, x


def free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def engine_values(tree, data, x, precision, fix_omega=False):
    e = engine.BranchSiteEngine(tree, data, precision=precision, fix_omega=fix_omega)
    try:
        nl = e.lnl(x)
        fh, site = e.fhK()
    finally:
        e.close()
    return nl, fh, site


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_equals_engine_two_calls(precision):
    print(f"case: noalloc {precision}: 40 taxa x 700 codons, then 9 taxa x 120 codons", flush=True)
    big = gene(1, 40, 700)
    small = gene(2, 9, 120, gaps=0.1)
    want = [engine_values(*g, precision) for g in (big, small)]
    got = [engine.lnl_noalloc(*g, precision=precision) for g in (big, small)]
    for (nl, fh, site), (nw, fw, sw) in zip(got, want):
        print(f"  -lnL {nl:.10f} engine {nw:.10f}", flush=True)
        assert nl == nw
        assert np.array_equal(fh, fw) and np.array_equal(site, sw)
    # H0 on the smaller gene again
    t, d, x = small
    nl0, _, _ = engine.lnl_noalloc(t, d, x[:-1], precision=precision, fix_omega=True)
    assert nl0 == engine_values(t, d, x[:-1], precision, fix_omega=True)[0]


def test_noalloc_allocates_nothing():
    print("case: noalloc device free memory over repeated calls", flush=True)
    g1, g2 = gene(3, 30, 400), gene(4, 12, 200)
    engine.lnl_noalloc(*g1)          # first call on the device: symbol address
    before = free_bytes()
    a = engine.lnl_noalloc(*g1)[0]
    b = engine.lnl_noalloc(*g2)[0]
    c = engine.lnl_noalloc(*g1)[0]
    after = free_bytes()
    print(f"  free before {before} after {after}; -lnL {a:.8f} {b:.8f} {c:.8f}", flush=True)
    assert after == before
    assert a == c
