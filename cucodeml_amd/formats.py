"""Input formats of the codeml likelihood path: Newick trees with ``#1``
branch labels, PHYLIP codon alignments, codon encoding with ambiguity codes,
site-pattern compression and F3x4 codon frequencies.

Node numbering follows codeml's tree reader (ReadTreeN, treesub.c:2883-3171):
tips are numbered 0..ns-1 in alignment order (matched by name), interior
nodes are numbered ns, ns+1, ... in the order their '(' appears, the first '('
is the root, and branches are numbered in the order nodes are attached, which
is the order of branch lengths in codeml's parameter vector (SetBranch,
treesub.c:3681-3693).
"""
from dataclasses import dataclass, field
import re
import numpy as np

from . import codons


@dataclass
class Tree:
    ns: int
    nnode: int
    root: int
    father: np.ndarray           # (nnode,) int, root -> -1
    sons: list                   # per node, ordered list of children
    branch: np.ndarray           # (nnode,) length of the edge above each node
    label: np.ndarray            # (nnode,) 0 background / 1 foreground, root -1
    ibranch: np.ndarray          # (nnode,) position in the branch-length vector
    names: list = field(default_factory=list)

    @property
    def nbranch(self):
        return self.nnode - 1

    def branch_vector(self):
        """Branch lengths in parameter-vector order."""
        t = np.zeros(self.nbranch)
        for v in range(self.nnode):
            if v != self.root:
                t[self.ibranch[v]] = self.branch[v]
        return t

    def set_branch_vector(self, t):
        for v in range(self.nnode):
            if v != self.root:
                self.branch[v] = t[self.ibranch[v]]

    def postorder(self):
        out, stack = [], [(self.root, False)]
        while stack:
            v, done = stack.pop()
            if done:
                out.append(v)
                continue
            stack.append((v, True))
            for s in reversed(self.sons[v]):
                stack.append((s, False))
        return out

    def newick(self, lengths=True, labels=True, names=None, digits=8):
        names = names or self.names or [f"S{i + 1}" for i in range(self.ns)]

        def rec(v):
            if v < self.ns:
                s = names[v]
            else:
                s = "(" + ",".join(rec(c) for c in self.sons[v]) + ")"
            if v != self.root:
                if labels and self.label[v] == 1:
                    s += " #1"
                if lengths:
                    s += f":{self.branch[v]:.{digits}f}"
            return s
        return rec(self.root) + ";"


def parse_newick(text, names):
    """Parse a Newick tree the way codeml's ReadTreeN numbers it.

    ``names`` is the list of sequence names; tips are matched to it.
    Branch labels ``#k`` and ``$k`` (clade label, pushed down to every
    branch of the clade) are supported, as are branch lengths.
    """
    ns = len(names)
    name_idx = {n: i for i, n in enumerate(names)}
    nmax = 2 * ns - 1
    father = np.full(nmax, -1, dtype=np.int64)
    sons = [[] for _ in range(nmax)]
    branch = np.zeros(nmax)
    label = np.full(nmax, -1.0)
    clade = np.full(nmax, -1, dtype=np.int64)
    ibranch = np.full(nmax, -1, dtype=np.int64)
    nnode, nbranch, root = ns, 0, -1
    cfather, inodeb, level = -1, 0, 0
    seen = np.zeros(ns, dtype=np.int64)
    i, L = 0, len(text)
    num = re.compile(r"\s*[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?")
    while i < L:
        ch = text[i]
        if ch == ";":
            if level != 0:
                raise ValueError("; inside parentheses in tree")
            break
        if ch == "," or not ch.isprintable() or ch.isspace():
            i += 1
            continue
        if ch == "(":
            level += 1
            cnode = nnode
            nnode += 1
            if nnode > nmax:
                raise ValueError("too many '(' for the number of sequences")
            if cfather >= 0:
                sons[cfather].append(cnode)
                father[cnode] = cfather
                ibranch[cnode] = nbranch
                nbranch += 1
            else:
                root = cnode
            cfather = cnode
            i += 1
        elif ch == ")":
            level -= 1
            inodeb = cfather
            cfather = father[cfather]
            i += 1
        elif ch in ":#$":
            m = num.match(text, i + 1)
            if not m:
                raise ValueError(f"number expected after {ch!r} at {i}")
            val = float(m.group(0))
            if ch == ":":
                branch[inodeb] = val
            elif ch == "#":
                label[inodeb] = val
            else:
                clade[inodeb] = int(val)
            i = m.end()
        else:
            if level <= 0:
                raise ValueError("expecting ; in the tree file")
            j = i
            while j < L and text[j] not in "(),:#$;" and text[j] != "\n":
                j += 1
            nm = text[i:j].strip()
            if nm not in name_idx:
                raise ValueError(f"species {nm} not in the alignment")
            c = name_idx[nm]
            father[c] = cfather
            sons[cfather].append(c)
            ibranch[c] = nbranch
            nbranch += 1
            inodeb = c
            seen[c] += 1
            i = j
    if np.any(seen != 1):
        bad = [names[k] for k in np.nonzero(seen != 1)[0]]
        raise ValueError(f"species missing or repeated in tree: {bad}")
    if nnode != nbranch + 1:
        raise ValueError("nnode != nbranch + 1")
    # clade labels '$k' apply to all branches below (DownTreeCladeLabel)
    if np.any(clade >= 0):
        def down(v, lab):
            if clade[v] >= 0:
                lab = clade[v]
            if v != root and lab >= 0:
                label[v] = lab
            for s in sons[v]:
                down(s, lab)
        down(root, -1)
    for v in range(nnode):
        if v != root and label[v] < 0:
            label[v] = 0
    label[root] = -1
    t = Tree(ns=ns, nnode=nnode, root=root, father=father[:nnode].copy(),
             sons=[list(s) for s in sons[:nnode]], branch=branch[:nnode].copy(),
             label=label[:nnode].astype(np.int64), ibranch=ibranch[:nnode].copy(),
             names=list(names))
    return t


# ---------------------------------------------------------------- alignments

def read_phylip(path_or_text):
    """Read a sequential PHYLIP nucleotide alignment (names then sequences)."""
    names, seqs = read_phylip_sets(path_or_text, 1)[0]
    return names, seqs


def read_phylip_sets(path_or_text, ndata):
    """The first `ndata` alignments of a file holding several sequential
    PHYLIP alignments one after another (codeml's ndata, codeml.c:315-343):
    a list of (names, seqs)."""
    if "\n" in path_or_text:
        text = path_or_text
    else:
        with open(path_or_text) as f:
            text = f.read()
    lines = text.splitlines()
    out, k = [], 0
    for i in range(ndata):
        while k < len(lines) and not lines[k].strip():
            k += 1
        if k >= len(lines):
            raise ValueError(f"the sequence file holds {i} data sets, ndata = {ndata}")
        names, seqs, k = _read_phylip_at(lines, k)
        out.append((names, seqs))
    return out


def _read_phylip_at(lines, k):
    """One alignment starting at line k (its "ns ls" line); returns (names,
    seqs, first line after it)."""
    ns, ls = (int(v) for v in lines[k].split()[:2])
    k += 1
    names, seqs = [], []
    for _ in range(ns):
        while not lines[k].strip():
            k += 1
        line = lines[k].rstrip()
        k += 1
        m = re.match(r"^(\S.*?)(\s{2,}|$)(.*)$", line)
        name, rest = m.group(1).strip(), m.group(3)
        buf = [c for c in rest if not c.isspace()]
        while len(buf) < ls:
            buf.extend(c for c in lines[k] if not c.isspace())
            k += 1
        names.append(name)
        seqs.append("".join(buf[:ls]).upper())
    return names, seqs, k


def write_phylip(path, names, seqs):
    with open(path, "w") as f:
        f.write(f"{len(names)} {len(seqs[0])}\n")
        for n, s in zip(names, seqs):
            f.write(f"{n}  {s}\n")


@dataclass
class CodonData:
    names: list
    ns: int
    ls: int                # codons
    npatt: int
    z: np.ndarray          # (ns, npatt) uint8: 0..60 sense, >=61 ambiguity code
    fpatt: np.ndarray      # (npatt,) float64 site-pattern counts
    amb_triplets: list     # triplet string of each ambiguity code 61..
    amb_sets: list         # sense codons each ambiguity code stands for
    pi: np.ndarray = None  # (61,) codon frequencies
    pose: np.ndarray = None  # (ls,) pattern index of every site

    @property
    def ncodes(self):
        return codons.NSENSE + len(self.amb_sets)

    def code_table(self):
        """(ncodes, 61) 0/1 matrix: row c marks the sense codons code c allows."""
        m = np.zeros((self.ncodes, codons.NSENSE), dtype=np.float64)
        m[np.arange(codons.NSENSE), np.arange(codons.NSENSE)] = 1.0
        for a, s in enumerate(self.amb_sets):
            m[codons.NSENSE + a, s] = 1.0
        return m


def code_amino_acid(data, c):
    """Amino acid of codon code c as codeml prints it (GetAASiteSpecies,
    codeml.c:4941-4971): '-' for a codon that allows every sense codon, '*'
    when the codons it allows code for more than one amino acid."""
    if c < codons.NSENSE:
        return str(codons.AA61[c])
    allowed = data.amb_sets[c - codons.NSENSE]
    if len(allowed) == codons.NSENSE:
        return "-"
    aas = {str(codons.AA61[k]) for k in allowed}
    return aas.pop() if len(aas) == 1 else "*"


def code_triplet(data, c):
    """Text of codon code c: the sense codon, or the ambiguous triplet as read
    (codeml's CODONs table, EncodeSeqs treesub.c:1097-1131)."""
    return codons.codon_str(c) if c < codons.NSENSE else data.amb_triplets[c - codons.NSENSE]


def sequences_of(data):
    """Nucleotide sequences of `data` site by site (each code's triplet text),
    e.g. to write the alignment back out as PHYLIP."""
    tri = [code_triplet(data, c) for c in range(data.ncodes)]
    cols = data.z[:, data.pose]
    return ["".join(tri[c] for c in row) for row in cols]


def site_amino_acid(data, species, site):
    """Amino acid of `species` at `site` as codeml prints it (code_amino_acid)."""
    return code_amino_acid(data, int(data.z[species, data.pose[site]]))


def encode_codons(names, seqs):
    """Turn nucleotide sequences into codon codes (EncodeSeqs, treesub.c:1075-1144).

    Fully resolved sense codons get 0..60; every distinct ambiguous triplet
    gets the next free code from 61 on; a resolved stop codon is an error.
    """
    ns = len(seqs)
    if any(len(s) % 3 for s in seqs):
        raise ValueError("sequence length is not a multiple of 3")
    ls = len(seqs[0]) // 3
    z = np.zeros((ns, ls), dtype=np.uint8)
    amb_index, amb_triplets, amb_sets = {}, [], []
    for j, s in enumerate(seqs):
        if len(s) != 3 * ls:
            raise ValueError("sequences differ in length")
        for h in range(ls):
            tri = s[3 * h:3 * h + 3]
            if all(c in codons.BASES for c in tri):
                c64 = 16 * codons.BASES.index(tri[0]) + 4 * codons.BASES.index(tri[1]) \
                    + codons.BASES.index(tri[2])
                k = codons.FROM64[c64]
                if k < 0:
                    raise ValueError(f"stop codon {tri} in seq {j + 1} ({names[j]})")
                z[j, h] = k
            else:
                if tri not in amb_index:
                    sset = codons.resolve_codon(tri)
                    if not sset:
                        raise ValueError(f"codon {tri} is a stop codon")
                    if codons.NSENSE + len(amb_triplets) > 255:
                        raise ValueError("too many distinct ambiguous codons")
                    amb_index[tri] = codons.NSENSE + len(amb_triplets)
                    amb_triplets.append(tri)
                    amb_sets.append(sset)
                z[j, h] = amb_index[tri]
    return z, amb_triplets, amb_sets


def compress_patterns(z, return_pose=False):
    """Collapse identical columns of z (ns, ls) into (ns, npatt) + counts
    (+ the pattern of every site, codeml's pose, with return_pose)."""
    cols, inv, counts = np.unique(z.T, axis=0, return_inverse=True, return_counts=True)
    out = (np.ascontiguousarray(cols.T), counts.astype(np.float64))
    return out + (inv.ravel().astype(np.int64),) if return_pose else out


def f3x4_frequencies(z, fpatt, amb_triplets, niter=20):
    """F3x4 codon frequencies with ambiguity resolution by iteration.

    Follows InitializeCodon / CountCodons / AddCodonFreqSeqGene
    (codeml.c:3511-3575, 3600-3660, 3663-3740): start from counts of fully
    resolved codons, then up to 20 rounds that share each ambiguous codon
    (and each ambiguous position) out in proportion to the current
    frequencies, stopping when codon, 3x4 and 1x4 tables all move by less
    than 1e-8 (Euclidean).  pi_i is the product of the three position
    frequencies, renormalised over the 61 sense codons.
    Every (species, pattern) cell contributes linearly, so cells are first
    summed per code.
    """
    ncode = codons.NSENSE + len(amb_triplets)
    w = np.bincount(z.ravel().astype(np.int64),
                    weights=np.broadcast_to(fpatt, z.shape).ravel(), minlength=ncode)
    trip = [codons.codon_str(i) for i in range(codons.NSENSE)] + list(amb_triplets)
    posb = [codons.position_bases(t) for t in trip]
    nb = [[len(p) for p in pb] for pb in posb]
    # initial counts: resolved codons only (nb product == 1)
    fcod = np.zeros(64)
    for c in range(ncode):
        if w[c] == 0:
            continue
        if nb[c][0] * nb[c][1] * nb[c][2] > 1:
            continue
        b = posb[c]
        fcod[16 * b[0][0] + 4 * b[1][0] + b[2][0]] += w[c]

    def fcodon_3x4(fc):
        t = np.zeros((3, 4))
        for k in range(64):
            t[0, k // 16] += fc[k]
            t[1, (k // 4) % 4] += fc[k]
            t[2, k % 4] += fc[k]
        return t / t.sum(axis=1, keepdims=True)

    f3 = fcodon_3x4(fcod)
    f1 = f3.sum(axis=0) / 3
    if len(amb_triplets):
        fc0, f30, f10 = fcod / fcod.sum(), f3.copy(), f1.copy()
        compat = []
        for c in range(ncode):
            b = posb[c]
            lst = []
            for b0 in b[0]:
                for b1 in b[1]:
                    for b2 in b[2]:
                        ic = 16 * b0 + 4 * b1 + b2
                        if codons.FROM64[ic] >= 0:
                            lst.append(ic)
            compat.append(lst)
        for _ in range(niter):
            fc = np.zeros(64)
            f3n = np.zeros((3, 4))
            f1n = np.zeros(4)
            for c in range(ncode):
                if w[c] == 0:
                    continue
                for k in range(3):
                    ib = posb[c][k]
                    t = sum(f30[k, b] for b in ib)
                    t1 = sum(f10[b] for b in ib)
                    for b in ib:
                        f3n[k, b] += w[c] * f30[k, b] / t
                        f1n[b] += w[c] * f10[b] / t1
                lst = compat[c]
                t = sum(fc0[ic] for ic in lst)
                for ic in lst:
                    fc[ic] += w[c] * fc0[ic] / t if t > 0 else w[c] / len(lst)
            fc /= fc.sum()
            f3n /= f3n.sum(axis=1, keepdims=True)
            f1n /= f1n.sum()
            d1 = np.sqrt(((fc - fc0) ** 2).sum())
            d2 = np.sqrt(((f3n - f30) ** 2).sum())
            d3 = np.sqrt(((f1n - f10) ** 2).sum())
            f3, f1 = f3n, f1n
            if d1 < 1e-8 and d2 < 1e-8 and d3 < 1e-8:
                break
            fc0, f30, f10 = fc, f3n, f1n
    c64 = codons.FROM61
    pi = f3[0, c64 // 16] * f3[1, (c64 // 4) % 4] * f3[2, c64 % 4]
    return pi / pi.sum()


def load_codon_data(names, seqs):
    z, trip, sets = encode_codons(names, seqs)
    zp, fpatt, pose = compress_patterns(z, return_pose=True)
    d = CodonData(names=list(names), ns=len(names), ls=z.shape[1], npatt=zp.shape[1],
                  z=zp, fpatt=fpatt, amb_triplets=trip, amb_sets=sets, pose=pose)
    d.pi = f3x4_frequencies(zp, fpatt, trip)
    return d
