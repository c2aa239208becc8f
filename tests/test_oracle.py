"""The CPU oracle (oracle/bsa_oracle.c) pinned against the reference's own CPU
codeml output (golden fixtures) and against mathematics that fixes each part:
rate-matrix identities, numpy/scipy eigen and matrix exponential, brute-force
sums over ancestral states, and probability sums over all site patterns."""
import itertools

import numpy as np
import pytest
import scipy.linalg

from cucodeml_amd import codons, formats, simulate
from oracle import oracle

from . import fixtures


@pytest.fixture(scope="module", params=fixtures.NAMES)
def golden(request):
    return fixtures.load(request.param)


def test_oracle_matches_reference_codeml(golden):
    fx, tree, data = golden
    L, _, site = oracle.lnl(tree, data, *fixtures.params(fx))
    # out prints lnL with %13.6f
    assert abs(L - fx["fixed_lnL"]) <= 6e-7 + 1e-12 * abs(L)
    assert abs(L - fx["fixed_lnL"]) / abs(L) < 1e-9
    ref = fixtures.site_to_pattern(fx, data)
    # lnf prints log f_h with %16.10f; sites whose f rests on tiny P(t)
    # entries (several substitutions on short branches) carry the rounding of
    # the cancelling sum U diag(e) V, so the bound is relative
    assert np.all(np.abs(site - ref) <= 1e-10 * np.abs(ref) + 1e-10)


def test_oracle_sum_of_sites_is_lnl(golden):
    fx, tree, data = golden
    L, fh, site = oracle.lnl(tree, data, *fixtures.params(fx))
    assert abs(np.dot(site, data.fpatt) - L) < 1e-9 * abs(L)
    fk, _, _, _, _ = oracle.model(*[fixtures.params(fx)[i] for i in (0, 1, 2, 3, 4)], data.pi)
    mix = np.log(np.exp(fh).T @ fk)
    assert np.max(np.abs(mix - site)) < 1e-10


def _pi(seed=3):
    return simulate.dirichlet_f3x4(np.random.default_rng(seed))[1]


def test_q_identities():
    pi = _pi()
    Q, mr = oracle.q_matrix(2.5, 0.3, pi)
    assert np.max(np.abs(Q.sum(1))) < 1e-14
    D = pi[:, None] * Q
    assert np.max(np.abs(D - D.T)) < 1e-15
    assert abs(mr + np.dot(pi, np.diag(Q))) < 1e-14
    ttt, ttc, tta, ctt = (int(codons.FROM64[c]) for c in (0, 1, 2, 16))
    assert Q[ttt, ttc] == pytest.approx(2.5 * pi[ttc], rel=1e-15)       # syn. transition
    assert Q[ttt, tta] == pytest.approx(0.3 * pi[tta], rel=1e-15)       # nonsyn. transversion
    assert Q[ttt, ctt] == pytest.approx(2.5 * 0.3 * pi[ctt], rel=1e-15)  # Phe->Leu transition
    assert Q[ttt, int(codons.FROM64[5])] == 0.0                         # two changes
    # independent numpy construction
    Qs, mrs = simulate.gy94_q(2.5, 0.3, pi)
    assert np.max(np.abs(Q - Qs)) < 1e-15 and abs(mr - mrs) < 1e-14


def test_eigen_against_numpy():
    pi = _pi(5)
    Q, _ = oracle.q_matrix(1.7, 2.2, pi)
    R, U, V = oracle.eigen(Q, pi)
    assert np.max(np.abs(U @ np.diag(R) @ V - Q)) < 1e-12
    assert np.max(np.abs(U @ V - np.eye(61))) < 1e-12
    s = np.sqrt(pi)
    lam = np.linalg.eigvalsh((s[:, None] * Q) / s[None, :])
    assert np.max(np.abs(np.sort(lam) - np.sort(R))) < 1e-12
    assert abs(R.max()) < 1e-12 and np.all(np.diff(R) <= 1e-15)


def test_eigen_with_zero_frequency_codons():
    pi = _pi(7)
    pi[[3, 17, 40]] = 0.0
    pi /= pi.sum()
    Q, _ = oracle.q_matrix(2.0, 0.5, pi)
    R, U, V = oracle.eigen(Q, pi)
    for t in (0.01, 0.3, 2.0):
        P = oracle.pmat(U, V, R, t)
        E = scipy.linalg.expm(Q * t)
        keep = pi > 0
        assert np.max(np.abs(P[np.ix_(keep, keep)] - E[np.ix_(keep, keep)])) < 1e-12
        assert np.allclose(P[~keep][:, ~keep], np.eye(3))


def test_pmat_expm_and_chapman_kolmogorov():
    pi = _pi(9)
    Q, _ = oracle.q_matrix(3.0, 0.05, pi)
    R, U, V = oracle.eigen(Q, pi)
    for t in (1e-4, 0.05, 0.5, 3.0):
        P = oracle.pmat(U, V, R, t)
        assert np.max(np.abs(P - np.clip(scipy.linalg.expm(Q * t), 0, None))) < 1e-12
        assert np.max(np.abs(P.sum(1) - 1)) < 1e-12
        D = pi[:, None] * P
        assert np.max(np.abs(D - D.T)) < 1e-13
    Ps, Pt, Pst = (oracle.pmat(U, V, R, t) for t in (0.2, 0.3, 0.5))
    assert np.max(np.abs(Ps @ Pt - Pst)) < 1e-12
    assert np.array_equal(oracle.pmat(U, V, R, 0.0), np.eye(61))


def _star_tree(ns, fg, bl):
    names = [f"S{i + 1}" for i in range(ns)]
    txt = "(" + ",".join(f"S{i + 1}{' #1' if i == fg else ''}:{bl[i]}" for i in range(ns)) + ");"
    return formats.parse_newick(txt, names), names


def _data(names, cols, amb=()):
    """CodonData from explicit code columns (list of tuples of triplets)."""
    seqs = ["".join(col[j] for col in cols) for j in range(len(names))]
    return formats.load_codon_data(names, seqs), seqs


def test_pruning_brute_force_four_taxa():
    rng = np.random.default_rng(1)
    names = ["S1", "S2", "S3", "S4"]
    txt = "((S1:0.12,S2 #1:0.3):0.07,S3:0.2,S4:0.05);"
    tree = formats.parse_newick(txt, names)
    pi = _pi(11)
    table = [codons.codon_str(i) for i in range(61)]
    cols = [tuple(table[int(k)] for k in rng.integers(61, size=4)) for _ in range(6)]
    cols.append(("---", "AAR", "NNN", "GCN"))
    data, _ = _data(names, cols)
    par = (2.2, 0.45, 0.35, 0.2, 4.0)
    L, fh, site = oracle.lnl(tree, data, *par, pi=pi)
    # P(t) from the oracle's eigen route (pinned against expm separately);
    # expm's own rounding of tiny entries would blur a 1e-11 check
    fk, qf, R3, U3, V3 = oracle.model(*par, pi)
    sets = [(0, 0), (1, 1), (0, 2), (1, 2)]

    def P(c, lab, t):
        s = sets[c][lab]
        return oracle.pmat(U3[s], V3[s], R3[s], t * qf[lab])
    props = fk
    mask = data.code_table()
    lab = {v: int(tree.label[v]) for v in range(tree.nnode)}
    a = tree.father[0]                      # interior node above S1, S2
    for h in range(data.npatt):
        x = [mask[data.z[s, h]] for s in range(4)]
        f = 0.0
        for c in range(4):
            tip = {s: P(c, lab[s], tree.branch[s]) @ x[s] for s in range(4)}
            inner = tip[0] * tip[1]
            up = P(c, lab[a], tree.branch[a]) @ inner
            fc = np.dot(pi, up * tip[2] * tip[3])
            assert abs(np.log(fc) - fh[c, h]) < 1e-12 * abs(fh[c, h])
            f += props[c] * fc
        assert abs(np.log(f) - site[h]) < 1e-12 * abs(site[h])
    # the class proportions and Qfactors used above are the model's own; the
    # independent numpy construction agrees with them
    props2, _, qf2 = simulate.model_a_classes(par[0], par[3], par[4], par[1], par[2], pi)
    assert np.max(np.abs(props2 - fk)) < 1e-15 and np.max(np.abs(np.array(qf2) - qf)) < 1e-12


def test_site_probabilities_sum_to_one():
    """Over all 61^2 patterns of two tips (third tip a gap) f sums to 1."""
    names = ["S1", "S2", "S3"]
    tree, _ = _star_tree(3, 1, [0.1, 0.4, 0.25])
    table = [codons.codon_str(i) for i in range(61)]
    cols = [(a, b, "---") for a, b in itertools.product(table, table)]
    data, _ = _data(names, cols)
    par = (1.8, 0.6, 0.3, 0.1, 2.5)
    _, fh, site = oracle.lnl(tree, data, *par, pi=_pi(12))
    assert abs(np.exp(site).sum() - 1) < 1e-12
    assert np.max(np.abs(np.exp(fh).sum(1) - 1)) < 1e-12
    # an all-gap column has probability 1
    d2, _ = _data(names, [("---", "---", "---"), ("NNN", "---", "?NN")])
    _, _, s2 = oracle.lnl(tree, d2, *par, pi=_pi(12))
    assert np.max(np.abs(s2)) < 1e-13


def test_scaling_does_not_change_lnl():
    fx, tree, data = fixtures.load("cfg1")
    assert oracle.scale_marks(tree)[0] > 0
    a, _, s1 = oracle.lnl(tree, data, *fixtures.params(fx), use_scaling=True)
    b, _, s2 = oracle.lnl(tree, data, *fixtures.params(fx), use_scaling=False)
    assert abs(a - b) < 1e-10 * abs(a)
    assert np.max(np.abs(s1 - s2)) < 1e-10


def test_class_limits():
    """p0 -> 1 leaves only class 0: the mixture equals that class."""
    fx, tree, data = fixtures.load("cfg0")
    k, _, _, w0, w2 = fixtures.params(fx)
    L, fh, site = oracle.lnl(tree, data, k, 1 - 1e-13, 0.5e-13, w0, w2)
    assert np.max(np.abs(site - fh[0])) < 1e-9
    L, fh, site = oracle.lnl(tree, data, k, 0.3, 0.7 - 1e-14, w0, w2)
    mix = np.log(0.3 * np.exp(fh[0]) + 0.7 * np.exp(fh[1]))
    assert np.max(np.abs(site - mix)) < 1e-9


@pytest.mark.parametrize("name", fixtures.IMPOSSIBLE)
def test_oracle_matches_reference_on_impossible_sites(name):
    """Sites no class can produce (sister tips on zero-length branches with
    different codons): the reference floors each class's f at 1e-300
    (fx_r treesub.c:7524-7534) and, where it scales nodes, adds the scale
    logs of NodeScale (treesub.c:6940-6968), -800 at a scaled node whose
    partials are all below 1e-300 (which it then resets to 1).  The oracle
    restates both and matches every site the reference prints; without
    node scaling those sites are log(1e-300) exactly."""
    fx, tree, data = fixtures.load(name)
    ref = fixtures.site_to_pattern(fx, data)
    L, _, site = oracle.lnl(tree, data, *fixtures.params(fx))
    assert abs(L - fx["fixed_lnL"]) <= 6e-7 + 1e-12 * abs(L)
    assert np.all(np.abs(site - ref) <= 1e-10 * np.abs(ref) + 1e-10)
    bad = np.isin(np.arange(data.npatt), fixtures.patterns_of_sites(fx, data,
                                                                    fx["impossible_sites"]))
    assert bad.sum() >= 1
    _, _, plain = oracle.lnl(tree, data, *fixtures.params(fx), use_scaling=False)
    assert np.all(np.abs(plain[bad] - np.log(1e-300)) < 1e-12)
    if name == "imposs8":
        assert np.all(np.abs(ref[bad] - np.log(1e-300)) < 1e-9)
    else:
        # the reference's scale logs make these sites ~-800 lower
        assert np.all(ref[bad] < np.log(1e-300) - 100)
