"""Host side of the codeml-compatible driver: control-file parsing and the
model-A checks (codeml.c:1430-1768), in.codeml reading (treesub.c:3925-3927),
the likelihood-ratio test's p-values against scipy's chi-square, and the out
file layout read back by the parser the tests use for the reference's own
out files (codeml.c:738-751)."""
import os
import tempfile

import numpy as np
import pytest
from scipy import stats

from cucodeml_amd import codeml
from . import fixtures


def write(d, name, text):
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def test_read_ctl_comments_and_values():
    with tempfile.TemporaryDirectory() as d:
        p = write(d, "c.ctl", "seqfile = a.phy * alignment\n* comment line\n  model = 2\n"
                              "omega = .4  * start\nfix_omega=1\n")
        c = codeml.read_ctl(p)
    assert c == {"seqfile": "a.phy", "model": "2", "omega": ".4", "fix_omega": "1"}


def test_check_ctl_accepts_model_a_rejects_others():
    with tempfile.TemporaryDirectory() as d:
        good = codeml.read_ctl(write(d, "g.ctl", fixtures.CTL.format(fix_omega=0, omega=1.5)))
    codeml.check_ctl(good)
    for k, v in [("model", "0"), ("NSsites", "1"), ("CodonFreq", "3"), ("clock", "1"),
                 ("cleandata", "1"), ("seqtype", "2"), ("fix_kappa", "1"), ("icode", "1"),
                 ("fix_blength", "-1"), ("fix_blength", "3")]:
        bad = dict(good, **{k: v})
        with pytest.raises(ValueError):
            codeml.check_ctl(bad)


def test_in_codeml_fixed_and_start():
    with tempfile.TemporaryDirectory() as d:
        fixed, v = codeml.read_in_codeml(write(d, "in.codeml", "-1 0.1 0.2\n2.0"))
        assert fixed and np.allclose(v, [0.1, 0.2, 2.0])
        fixed, v = codeml.read_in_codeml(write(d, "in2", "0.1 0.2 2.0\n"))
        assert not fixed and np.allclose(v, [0.1, 0.2, 2.0])


@pytest.mark.parametrize("d", [0.0, 0.3, 2.71, 3.84, 10.0, 40.0])
def test_lrt_p_values(d):
    h0 = codeml.Fit("H0", -100.0, np.zeros(3), 2, .5, .3, .1, 1, np.zeros(4), np.zeros(1), 3, 0,
                    0, True, 0)
    h1 = codeml.Fit("H1", -100.0 + d / 2, np.zeros(3), 2, .5, .3, .1, 2, np.zeros(4),
                    np.zeros(1), 4, 0, 0, True, 0)
    r = codeml.lrt(h0, h1)
    assert abs(r["stat"] - d) < 1e-12
    assert abs(r["p_chi2_1"] - stats.chi2.sf(d, 1)) < 1e-12
    assert abs(r["p_mixture"] - 0.5 * stats.chi2.sf(d, 1)) < 1e-12


def test_out_file_layout_round_trip():
    fx, tree, data = fixtures.load("cfg0")
    nb = tree.nbranch
    x = np.concatenate([np.linspace(0.01, 0.3, nb), [1.94, np.log(0.5 / 0.2), np.log(0.3 / 0.2),
                                                     0.05, 4.2]])
    fit = codeml.Fit("H1", -2241.441189, x, 1.94, 0.5, 0.3, 0.05, 4.2,
                     np.array([0.5, 0.3, 0.12, 0.08]), x[:nb], len(x), 10, 12, True, 0.5)
    text = codeml.format_out(tree, data, fit, fx["names"])
    lnl, npar, xr = fixtures.parse_out(text)
    assert lnl == -2241.441189 and npar == len(x)
    want = fit.x_untransformed()
    assert np.allclose(xr, np.round(want, 6), atol=1e-12)
    assert abs(xr[nb + 1] - 0.5) < 1e-12 and abs(xr[nb + 2] - 0.3) < 1e-12


def test_load_problems_reads_ndata_sets_on_one_tree():
    fx, tree, data = fixtures.load("cfg0")
    lines = fx["phylip"].strip().splitlines()
    head, rows = lines[0], [ln for ln in lines[1:] if ln.strip()]
    # the second data set lists the same taxa in reverse order
    second = head + "\n" + "\n".join(reversed(rows)) + "\n"
    with tempfile.TemporaryDirectory() as d:
        write(d, "seq.phy", fx["phylip"] + "\n" + second)
        write(d, "tree.nwk", f"{fx['ns']} 1\n{fx['newick']}\n")
        ctl = codeml.read_ctl(write(d, "codeml.ctl", fixtures.CTL.format(fix_omega=0, omega=1.5)
                                    + "ndata = 2\n"))
        probs = codeml.load_problems(ctl, d)
        assert len(probs) == 2
        ctl1 = dict(ctl, ndata="1")
        assert len(codeml.load_problems(ctl1, d)) == 1
        ctl3 = dict(ctl, ndata="3")
        with pytest.raises(ValueError):
            codeml.load_problems(ctl3, d)
    (t1, d1, n1), (t2, d2, n2) = probs
    assert n2 == list(reversed(n1))
    assert d1.npatt == data.npatt and d2.npatt == data.npatt
    # the tree numbers tips in the order of each alignment, so the two
    # problems are the same up to that relabelling: same branch lengths
    # as a multiset and the same compressed site counts
    assert np.allclose(np.sort(t1.branch_vector()), np.sort(t2.branch_vector()))
    assert np.array_equal(np.sort(d1.fpatt), np.sort(d2.fpatt))


def test_fix_blength_2_needs_branch_lengths():
    fx, _, _ = fixtures.load("cfg0")
    with tempfile.TemporaryDirectory() as d:
        write(d, "seq.phy", fx["phylip"])
        from cucodeml_amd import formats
        names, _ = formats.read_phylip(fx["phylip"])
        t = formats.parse_newick(fx["newick"], names)
        write(d, "tree.nwk", t.newick(lengths=False))
        ctl = fixtures.CTL.format(fix_omega=0, omega=1.5).replace("fix_blength = 0",
                                                                  "fix_blength = 2")
        c = codeml.read_ctl(write(d, "codeml.ctl", ctl))
        codeml.check_ctl(c)
        with pytest.raises(ValueError, match="no branch lengths"):
            codeml.load_problems(c, d)
        c["fix_blength"] = "1"
        assert len(codeml.load_problems(c, d)) == 1


def _lnf_rows(text):
    rows = [ln.split() for ln in text.splitlines()]
    return rows[0], [r for r in rows[1:] if len(r) > 5]


def test_lnf_file_matches_reference_layout_and_order():
    """The driver's lnf text (format_lnf) from the oracle's per-site values
    against the reference's own lnf file for the same data and parameters
    (ml16 fixed-parameter run): header, pattern order (PatternWeight's
    sorted column text, treesub.c:1302-1435), counts and codon/amino-acid
    columns exactly, the numbers to their print precision."""
    from oracle import oracle
    fx, tree, data = fixtures.load("ml16")
    _, _, site = oracle.lnl(tree, data, *fixtures.params(fx))
    got = codeml.format_lnf(data, site)
    ref = fx["fixed_lnf_text"]
    hg, rg = _lnf_rows(got)
    hr, rr = _lnf_rows(ref)
    assert hg == hr and len(rg) == len(rr) == data.npatt
    assert got.count("\n") == ref.count("\n")
    for a, b in zip(rg, rr):
        assert a[0] == b[0] and a[1] == b[1] and a[5:] == b[5:]
        assert abs(float(a[2]) - float(b[2])) <= 1e-10 * abs(float(b[2])) + 2e-10
        assert abs(float(a[3]) - float(b[3])) <= 2e-12 + 1e-9 * float(b[3])
        assert abs(float(a[4]) - float(b[4])) <= 1.5e-4
