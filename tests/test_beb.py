"""Site posteriors on the host: the oracle's restatement of the naive and
Bayes empirical Bayes analyses (oracle.neb / beb_sets / beb_integrate,
following codeml.c:5079-5129 and 6419-6826) against what the reference's CPU
codeml prints for the same data at the same parameters (golden fixtures from
tests/golden/make_beb_golden.py: per-site posteriors at %5.5f, the grid
posteriors, log f(X)); the site-to-pattern map, the first sequence's amino
acids, and the out/rst layouts read back."""
import gzip
import json
import os

import numpy as np
import pytest

from cucodeml_amd import codeml, formats
from oracle import oracle
from . import fixtures

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_beb(name):
    with gzip.open(os.path.join(GOLDEN, f"beb_{name}.json.gz"), "rt") as f:
        return json.load(f)


def test_site_pattern_map_and_amino_acids():
    for name in ("toy_amb", "cfg0", "cfg4s"):
        fx, _, data = fixtures.load(name)
        gold = load_beb(name)
        z, _, _ = formats.encode_codons(fx["names"], fx["seqs"])
        assert np.array_equal(data.z[:, data.pose], z)
        assert np.array_equal(np.bincount(data.pose, minlength=data.npatt), data.fpatt)
        aa = "".join(formats.site_amino_acid(data, 0, i) for i in range(data.ls))
        assert aa == gold["aa"]


def test_oracle_neb_and_beb_match_reference():
    fx, tree, data = fixtures.load("toy_amb")
    gold = load_beb("toy_amb")
    kappa, p0, p1, w0, w2 = fixtures.params(fx)
    neb = oracle.neb(tree, data, kappa, p0, p1, w0, w2)
    F = oracle.beb_sets(tree, data, kappa, p0, p1, w0, w2)
    post, lfx, pw, p01 = oracle.beb_integrate(F, data.fpatt)
    # printed with 5 decimals: |d| <= 5e-6 plus the oracle's own rounding
    assert np.abs(neb[:, data.pose].T - np.array(gold["neb"])).max() <= 5.01e-6
    assert np.abs(post[:, data.pose].T - np.array(gold["beb"])).max() <= 5.01e-6
    assert abs(lfx - gold["log_fX"]) <= 1e-6
    assert np.allclose(np.round(pw[0], 3), gold["grid"]["post_w0"], atol=1e-12)
    assert np.allclose(np.round(pw[1], 3), gold["grid"]["post_w2"], atol=1e-12)
    assert np.allclose(np.round(p01, 3), gold["post_p0p1"], atol=1e-12)
    # posteriors are distributions over the four classes
    assert np.allclose(post.sum(axis=0), 1, atol=1e-12)
    assert np.allclose(neb.sum(axis=0), 1, atol=1e-12)


def test_grid_matches_reference_printout():
    gold = load_beb("cfg0")
    pc, iw, _ = oracle.beb_grid()
    assert np.allclose((np.arange(10) + 0.5) / 10, gold["grid"]["w0"])
    assert np.allclose(1 + (np.arange(10) + 0.5), gold["grid"]["w2"])
    # each grid point's class proportions are a distribution; the p0-p1
    # triangles cover the simplex with equal weight (centroids average to 1/3)
    assert np.allclose(pc.sum(axis=1), 1)
    q0, q1 = pc[::100, 0], pc[::100, 1]
    assert np.allclose([q0.mean(), q1.mean()], 1 / 3)
    assert iw.min() == 0 and iw.max() == 120


def test_out_and_rst_layout_round_trip():
    fx, _, data = fixtures.load("cfg0")
    gold = load_beb("cfg0")
    post = np.zeros((4, data.npatt))
    post[:, data.pose] = np.array(gold["beb"]).T
    beb = {"post": post, "post_w0": np.array(gold["grid"]["post_w0"]),
           "post_w2": np.array(gold["grid"]["post_w2"]),
           "post_p0p1": np.array(gold["post_p0p1"]), "log_fX": gold["log_fX"]}
    text = codeml.format_beb_out(beb, data)
    got = codeml.positive_sites(post, data)
    # the fixture's per-class values carry 5 decimals, so their sum may round
    # to the neighbouring third decimal
    assert [(s, a) for s, a, _, _ in got] == [(s, a) for s, a, _, _ in gold["positive_sites"]]
    assert np.allclose([t for _, _, t, _ in got], [t for _, _, t, _ in gold["positive_sites"]],
                       atol=1.01e-3)
    assert all(f"{s:6d} {a} {t:.3f}" in text for s, a, t, _ in got)
    rst = codeml.format_site_table("BEB", post, data, fx["names"])
    first = rst.splitlines()[3].split()
    assert first[0] == "1" and first[1] == gold["aa"][0]
    assert np.allclose([float(v) for v in first[2:6]], gold["beb"][0], atol=1e-12)
