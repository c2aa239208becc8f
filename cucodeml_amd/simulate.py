"""Synthetic data for the branch-site model A likelihood path.

Random Yule trees (unrooted, one ``#1`` foreground branch), F3x4 codon
frequencies with Dirichlet(1) per codon position, and codon alignments
simulated under GY94 / branch-site model A with branch lengths scaled the
way the likelihood scales them (Qfactor per branch type, codeml.c:2490-2512).
Everything is seeded; nothing here is on the timed path.
"""
import numpy as np

from . import codons
from .formats import Tree

_NDIFF, _POS, _TS, _SYN = codons.diff_tables()


def gy94_q(kappa, omega, pi):
    """Unscaled GY94 rate matrix (eigenQcodon, codeml.c:3147-3186) and its mean rate."""
    q = np.where(_NDIFF == 1, 1.0, 0.0)
    q = q * np.where(_TS, kappa, 1.0) * np.where(_SYN, 1.0, omega)
    Q = q * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mr = -np.dot(pi, np.diag(Q))
    return Q, mr


def model_a_classes(kappa, w0, w2, p0, p1, pi):
    """Class proportions, per-class (background, foreground) omegas and Qfactors."""
    t = p0 + p1
    props = np.array([p0, p1, (1 - t) * p0 / t, (1 - t) * p1 / t])
    omegas = [(w0, w0), (1.0, 1.0), (w0, w2), (1.0, w2)]
    _, mr_b = gy94_q(kappa, (p0 * w0 + p1) / t, pi)
    _, mr_f = gy94_q(kappa, p0 * w0 + p1 + (1 - t) * w2, pi)
    return props, omegas, (1.0 / mr_b, 1.0 / mr_f)


def _eig_rev(Q, pi):
    s = np.sqrt(pi)
    S = (s[:, None] * Q) / s[None, :]
    S = 0.5 * (S + S.T)
    lam, R = np.linalg.eigh(S)
    return lam, R, s


def _expm_rev(Q, pi, t, eig=None):
    lam, R, s = _eig_rev(Q, pi) if eig is None else eig
    P = (R * np.exp(lam * t)[None, :]) @ R.T
    P = P / s[:, None] * s[None, :]
    P = np.clip(P, 0.0, None)
    return P / P.sum(axis=1, keepdims=True)


def dirichlet_f3x4(rng):
    f3 = rng.dirichlet(np.ones(4), size=3)
    c = codons.FROM61
    pi = f3[0, c // 16] * f3[1, (c // 4) % 4] * f3[2, c % 4]
    return f3, pi / pi.sum()


def _finish_tree(ns, father, bl, rng, names):
    """Numbered tree (codeml order) from a rooted binary parent array, unrooted."""
    nn = len(father)
    ch = [[] for _ in range(nn)]
    root = int(np.nonzero(father < 0)[0][0])
    for v in range(nn):
        if father[v] >= 0:
            ch[father[v]].append(v)
    # unroot: merge an interior child of a bifurcating root into the root
    if len(ch[root]) == 2:
        a, b = ch[root]
        keep, drop = (a, b) if len(ch[b]) else (b, a)
        if len(ch[drop]) == 0:
            raise ValueError("cannot unroot a 2-tip tree")
        bl[keep] += bl[drop]
        ch[root] = [keep] + ch[drop]
        for c in ch[drop]:
            father[c] = root
        ch[drop] = []
        father[drop] = -2
    # renumber in preorder: tips keep 0..ns-1, interior by first visit
    order, stack = [], [root]
    while stack:
        v = stack.pop()
        order.append(v)
        stack.extend(reversed(ch[v]))
    new = {}
    nxt = ns
    for v in order:
        if len(ch[v]) == 0:
            new[v] = v
        else:
            new[v] = nxt
            nxt += 1
    nnode = nxt
    fa = np.full(nnode, -1, dtype=np.int64)
    sons = [[] for _ in range(nnode)]
    branch = np.zeros(nnode)
    ibranch = np.full(nnode, -1, dtype=np.int64)
    nb = 0
    for v in order:
        if v != root:
            ibranch[new[v]] = nb
            nb += 1
            fa[new[v]] = new[father[v]]
            branch[new[v]] = bl[v]
        sons[new[v]] = [new[c] for c in ch[v]]
    label = np.zeros(nnode, dtype=np.int64)
    rootn = new[root]
    label[rootn] = -1
    fg = int(rng.choice([v for v in range(nnode) if v != rootn]))
    label[fg] = 1
    return Tree(ns=ns, nnode=nnode, root=rootn, father=fa, sons=sons, branch=branch,
                label=label, ibranch=ibranch, names=list(names))


def yule_tree(ns, mean_bl, rng, names=None):
    """Random Yule topology, iid Exp(mean_bl) branch lengths, unrooted, one #1."""
    names = names or [f"S{i + 1}" for i in range(ns)]
    # grow by splitting a uniformly chosen leaf; leaves get tip ids at the end
    father = [-1]
    leaves = [0]
    while len(leaves) < ns:
        k = int(rng.integers(len(leaves)))
        v = leaves.pop(k)
        for _ in range(2):
            father.append(v)
            leaves.append(len(father) - 1)
    nn = len(father)
    # relabel: tips first (0..ns-1, random order), interior after
    tips = list(rng.permutation(leaves))
    interior = [v for v in range(nn) if v not in set(leaves)]
    mapping = {v: i for i, v in enumerate(tips)}
    mapping.update({v: ns + i for i, v in enumerate(interior)})
    fa = np.full(nn, -1, dtype=np.int64)
    for v in range(nn):
        if father[v] >= 0:
            fa[mapping[v]] = mapping[father[v]]
    bl = rng.exponential(mean_bl, size=nn)
    return _finish_tree(ns, fa, bl, rng, names)


def balanced_tree(ns, mean_bl, rng, names=None):
    """Fully balanced binary topology on ns (a power of two) tips, unrooted."""
    names = names or [f"S{i + 1}" for i in range(ns)]
    father = []
    level = list(range(ns))
    nxt = ns
    fa = {}
    while len(level) > 1:
        nl = []
        for a, b in zip(level[0::2], level[1::2]):
            fa[a] = nxt
            fa[b] = nxt
            nl.append(nxt)
            nxt += 1
        level = nl
    father = np.full(nxt, -1, dtype=np.int64)
    for k, v in fa.items():
        father[k] = v
    bl = rng.exponential(mean_bl, size=nxt)
    return _finish_tree(ns, father, bl, rng, names)


def simulate_model_a(tree, ncodon, kappa, w0, w2, p0, p1, pi, rng, gap_frac=0.0):
    """Codon alignment (list of nucleotide strings) under branch-site model A."""
    props, omegas, qf = model_a_classes(kappa, w0, w2, p0, p1, pi)
    cls = rng.choice(4, size=ncodon, p=props / props.sum())
    Qs, eig = {}, {}
    for w in {w0, 1.0, w2}:
        Qs[w], _ = gy94_q(kappa, w, pi)
        eig[w] = _eig_rev(Qs[w], pi)
    states = np.zeros((tree.nnode, ncodon), dtype=np.int64)
    cum_pi = np.cumsum(pi)
    states[tree.root] = np.minimum(np.searchsorted(cum_pi, rng.random(ncodon) * cum_pi[-1]), 60)
    order = []
    stack = [tree.root]
    while stack:
        v = stack.pop()
        order.append(v)
        stack.extend(tree.sons[v])
    for v in order:
        if v == tree.root:
            continue
        lab = int(tree.label[v])
        par = states[tree.father[v]]
        for c in range(4):
            idx = np.nonzero(cls == c)[0]
            if idx.size == 0:
                continue
            w = omegas[c][lab]
            P = _expm_rev(Qs[w], pi, tree.branch[v] * qf[lab], eig[w])
            cum = np.cumsum(P, axis=1)
            u = rng.random(idx.size) * cum[par[idx], -1]
            row = cum[par[idx]]
            states[v, idx] = np.minimum((row < u[:, None]).sum(axis=1), 60)
    seqs = []
    table = [codons.codon_str(i) for i in range(codons.NSENSE)]
    for j in range(tree.ns):
        s = [table[k] for k in states[j]]
        if gap_frac > 0:
            for h in np.nonzero(rng.random(ncodon) < gap_frac)[0]:
                s[h] = "---"
        seqs.append("".join(s))
    return seqs, cls


def uniform_params(rng):
    """kappa ~ U(1,4), w0 ~ U(0.01,0.5), w2 ~ U(1.5,6); p0, p1 = 0.5, 0.3."""
    return dict(kappa=float(rng.uniform(1, 4)), w0=float(rng.uniform(0.01, 0.5)),
                w2=float(rng.uniform(1.5, 6)), p0=0.5, p1=0.3)
