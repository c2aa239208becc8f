"""Input formats: Newick numbering as codeml's ReadTreeN, codon encoding,
pattern compression, F3x4 frequencies; the synthetic-data generator."""
import numpy as np
import pytest

from cucodeml_amd import codons, formats, simulate


def test_codon_tables():
    assert codons.NSENSE == 61
    assert [codons.codon_str(i) for i in (0, 1, 60)] == ["TTT", "TTC", "GGG"]
    assert codons.resolve_codon("---") == list(range(61))
    assert codons.resolve_codon("TAR") == []                    # TAA, TAG stops
    assert [codons.codon_str(i) for i in codons.resolve_codon("TTY")] == ["TTT", "TTC"]


def test_newick_numbering_follows_readtreen():
    names = ["a", "b", "c", "d", "e"]
    t = formats.parse_newick("((a:0.1,b:0.2):0.3,(c:0.4,d #1:0.5):0.6,e:0.7);", names)
    assert t.root == 5 and t.nnode == 8
    assert list(t.father) == [6, 6, 7, 7, 5, -1, 5, 5]
    assert t.sons[5] == [6, 7, 4] and t.sons[6] == [0, 1] and t.sons[7] == [2, 3]
    # branches in attachment order: node6, a, b, node7, c, d, e
    assert [int(t.ibranch[v]) for v in (6, 0, 1, 7, 2, 3, 4)] == list(range(7))
    assert np.allclose(t.branch_vector(), [0.3, 0.1, 0.2, 0.6, 0.4, 0.5, 0.7])
    assert list(t.label) == [0, 0, 0, 1, 0, -1, 0, 0]
    t2 = formats.parse_newick(t.newick(), names)
    assert np.array_equal(t2.father, t.father) and np.allclose(t2.branch, t.branch)
    with pytest.raises(ValueError):
        formats.parse_newick("((a,b),c,d);", names)


def test_clade_labels_push_down():
    names = ["a", "b", "c", "d"]
    t = formats.parse_newick("((a,b)$1,c,d);", names)
    assert t.label[t.father[0]] == 1 and t.label[0] == 1 and t.label[1] == 1
    assert t.label[2] == 0


def test_encode_and_compress():
    names = ["x", "y", "z"]
    seqs = ["TTTAARTTT---", "TTCAAATTCGGG", "TTTAAGTTTGGG"]
    z, trip, sets = formats.encode_codons(names, seqs)
    assert trip == ["AAR", "---"]
    assert z[0, 1] == 61 and z[0, 3] == 62
    zp, fp = formats.compress_patterns(z)
    assert zp.shape == (3, 3) and fp.sum() == 4 and sorted(fp) == [1, 1, 2]
    with pytest.raises(ValueError):
        formats.encode_codons(["x"], ["TAA"])


def test_f3x4_clean_data_is_position_product():
    rng = np.random.default_rng(0)
    tab = [codons.codon_str(i) for i in range(61)]
    seqs = ["".join(tab[k] for k in rng.integers(61, size=200)) for _ in range(5)]
    d = formats.load_codon_data([f"s{i}" for i in range(5)], seqs)
    cnt = np.zeros((3, 4))
    for s in seqs:
        for h in range(0, len(s), 3):
            for p in range(3):
                cnt[p, "TCAG".index(s[h + p])] += 1
    f = cnt / cnt.sum(1, keepdims=True)
    c = codons.FROM61
    pi = f[0, c // 16] * f[1, (c // 4) % 4] * f[2, c % 4]
    assert np.max(np.abs(d.pi - pi / pi.sum())) < 1e-15


def test_phylip_roundtrip(tmp_path):
    p = tmp_path / "a.phy"
    formats.write_phylip(str(p), ["s1", "long name"], ["AAATTT", "CCCGGG"])
    names, seqs = formats.read_phylip(str(p))
    assert names == ["s1", "long name"] and seqs == ["AAATTT", "CCCGGG"]


def test_phylip_several_data_sets(tmp_path):
    # ndata alignments one after another in one file (codeml.c:315-343)
    p = tmp_path / "m.phy"
    text = ("2 6\ns1  AAATTT\ns2  CCCGGG\n\n\n"
            "3 3\nb  TTT\na  GGG\nc  AAA\n"
            "2 9\nx  AAA CCC\nGGG\ny  TTTTTTTTT\n")
    p.write_text(text)
    sets = formats.read_phylip_sets(str(p), 3)
    assert sets[0] == (["s1", "s2"], ["AAATTT", "CCCGGG"])
    assert sets[1] == (["b", "a", "c"], ["TTT", "GGG", "AAA"])
    assert sets[2] == (["x", "y"], ["AAACCCGGG", "TTTTTTTTT"])
    assert formats.read_phylip_sets(str(p), 1) == sets[:1]
    assert formats.read_phylip(str(p)) == sets[0]
    with pytest.raises(ValueError, match="3 data sets"):
        formats.read_phylip_sets(str(p), 4)


@pytest.mark.parametrize("ns", [4, 12, 64])
def test_yule_tree_shape(ns):
    t = simulate.yule_tree(ns, 0.05, np.random.default_rng(ns))
    assert t.nnode == 2 * ns - 2 and len(t.sons[t.root]) == 3
    assert sorted(t.ibranch[t.ibranch >= 0]) == list(range(2 * ns - 3))
    assert (t.label == 1).sum() == 1
    t2 = formats.parse_newick(t.newick(digits=12), t.names)
    assert np.array_equal(t2.father, t.father)
    assert np.allclose(t2.branch_vector(), t.branch_vector())


def test_simulated_alignment_is_model_a_like():
    rng = np.random.default_rng(2)
    t = simulate.balanced_tree(8, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, cls = simulate.simulate_model_a(t, 300, 2.0, 0.1, 3.0, 0.5, 0.3, pi, rng, gap_frac=0.1)
    assert len(seqs) == 8 and all(len(s) == 900 for s in seqs)
    frac = np.bincount(cls, minlength=4) / 300
    assert abs(frac[0] - 0.5) < 0.1 and abs(frac[1] - 0.3) < 0.1
    gaps = sum(s.count("---") for s in seqs) / (8 * 300)
    assert 0.05 < gaps < 0.15
