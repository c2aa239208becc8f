"""The C ABI library: it loads, exports every entry point include/bsa.h
declares, and rejects bad problems with a status code and a message (these
checks run before any HIP call, so they need no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "bsa.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(bsa_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    names = header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), f"{n} declared in bsa.h but not exported"
    assert sorted(engine.EXPORTED) == names
    assert b"gfx950" in L.bsa_version()


def _problem(tree, data, **over):
    keep = []

    def ia(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int))
    z = np.ascontiguousarray(data.z, dtype=np.uint8)
    m = np.ascontiguousarray(data.code_table().astype(np.uint8))
    fp = np.ascontiguousarray(data.fpatt)
    pi = np.ascontiguousarray(data.pi)
    keep += [z, m, fp, pi]
    kw = dict(ns=tree.ns, nnode=tree.nnode, root=tree.root, father=ia(tree.father),
              label=ia(np.maximum(tree.label, 0)), ibranch=ia(tree.ibranch), npatt=data.npatt,
              z=z.ctypes.data_as(C.POINTER(C.c_ubyte)), ncodes=data.ncodes,
              code_mask=m.ctypes.data_as(C.POINTER(C.c_ubyte)),
              fpatt=fp.ctypes.data_as(C.POINTER(C.c_double)),
              pi=pi.ctypes.data_as(C.POINTER(C.c_double)), precision=0, kernel=0, device=0,
              tile=0)
    for k, v in over.items():
        kw[k] = ia(v) if isinstance(v, np.ndarray) else v
    return engine._Problem(**kw), keep


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(4)
    tree = simulate.yule_tree(6, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, 40, 2, 0.1, 3, 0.5, 0.3, pi, rng)
    return tree, formats.load_codon_data(tree.names, seqs)


@pytest.mark.parametrize("change, msg", [
    (dict(ns=2), "at least 3"),
    (dict(root=0), "interior"),
    (dict(npatt=0), "npatt"),
    (dict(precision=7), "precision"),
    (dict(kernel=5), "kernel"),
    (dict(kernel=2), "FP32 partials"),     # BSA_KERNEL_MFMA32 with FP64 partials
    (dict(kernel=3), "FP32 partials"),     # BSA_KERNEL_FP32MC with FP64 partials
    (dict(ncodes=300), "ncodes"),
    (dict(tile=96), "tile"),
])
def test_create_rejects_bad_problems(small, change, msg):
    tree, data = small
    prob, keep = _problem(tree, data, **change)
    h = C.c_void_p()
    rc = engine.lib().bsa_create(C.byref(prob), C.byref(h))
    assert rc == 1 and not h.value
    assert msg in engine.lib().bsa_last_error().decode()


def test_create_rejects_bad_trees(small):
    tree, data = small
    lab = np.maximum(tree.label, 0).copy()
    lab[lab == 1] = 0                       # no foreground branch
    prob, keep = _problem(tree, data, label=lab)
    h = C.c_void_p()
    assert engine.lib().bsa_create(C.byref(prob), C.byref(h)) == 1
    assert "foreground" in engine.lib().bsa_last_error().decode()
    lab = np.maximum(tree.label, 0).copy()
    lab[0] = 2                              # model A has two branch types
    prob, keep = _problem(tree, data, label=lab)
    assert engine.lib().bsa_create(C.byref(prob), C.byref(h)) == 1
    ib = tree.ibranch.copy()
    ib[0] = ib[1]                           # not a permutation
    prob, keep = _problem(tree, data, ibranch=ib)
    assert engine.lib().bsa_create(C.byref(prob), C.byref(h)) == 1
    assert "permutation" in engine.lib().bsa_last_error().decode()
    fa = tree.father.copy()
    fa[0] = 1                               # a tip as father
    prob, keep = _problem(tree, data, father=fa)
    assert engine.lib().bsa_create(C.byref(prob), C.byref(h)) == 1


def test_null_arguments_are_errors():
    L = engine.lib()
    assert L.bsa_create(None, None) == 1
    assert L.bsa_lnl(None, None, 0, None) == 1
    assert L.bsa_destroy(None) == 0
    assert L.bsa_num_params(None) == -1


def test_x_from_params_roundtrip():
    x = engine.x_from_params([0.1, 0.2, 0.3], 2.0, 0.5, 0.3, 0.1, 3.0)
    a, b = np.exp(x[4]), np.exp(x[5])
    assert np.allclose([a / (1 + a + b), b / (1 + a + b)], [0.5, 0.3])
    assert len(engine.x_from_params([0.1] * 3, 2.0, 0.5, 0.3, 0.1)) == 7


def test_host_model_setup_matches_oracle():
    """The product's host eigen route (Householder + QL) against the oracle's
    (Jacobi): proportions, Qfactors, eigenvalues and P(t) = U e^{Rt} V."""
    from oracle import oracle
    rng = np.random.default_rng(8)
    _, pi = simulate.dirichlet_f3x4(rng)
    par = (2.7, 0.55, 0.25, 0.07, 4.5)
    fk, qf, R, U, V = engine.model_a_setup(*par, pi)
    fko, qfo, Ro, Uo, Vo = oracle.model(*par, pi)
    assert np.allclose(fk, fko, rtol=1e-15) and np.allclose(qf, qfo, rtol=1e-14)
    for s in range(3):
        assert np.max(np.abs(np.sort(R[s]) - np.sort(Ro[s]))) < 1e-13
        assert np.max(np.abs(U[s] @ V[s] - np.eye(61))) < 1e-12
        for t in (0.01, 0.2, 1.5):
            P = (U[s] * np.exp(R[s] * t)) @ V[s]
            Po = oracle.pmat(Uo[s], Vo[s], Ro[s], t)
            assert np.max(np.abs(np.clip(P, 0, None) - Po)) < 1e-13
    # a codon with zero frequency is split off with root 0 and identity rows
    pi0 = pi.copy()
    pi0[[5, 30]] = 0
    pi0 /= pi0.sum()
    fk, qf, R, U, V = engine.model_a_setup(*par, pi0)
    Q, _ = oracle.q_matrix(par[0], par[3], pi0)
    keep = pi0 > 0
    QQ = U[0] @ np.diag(R[0]) @ V[0]
    assert np.max(np.abs((QQ - Q)[np.ix_(keep, keep)])) < 1e-12
    assert U[0][5, 5] == 1 and V[0][30, 30] == 1 and R[0][5] == 0


@pytest.mark.parametrize("kind", ["uniform", "two", "three", "five"])
def test_host_eigen_degenerate_and_small_systems(kind):
    """Equal codon frequencies give repeated eigenvalues (the genetic code's
    symmetries), and frequencies on only 2, 3 or 5 codons leave tiny
    systems: the reconstruction Q = U diag(R) V, U V = I and the eigenvalues
    against numpy's symmetric solver hold there too."""
    from oracle import oracle
    if kind == "uniform":
        pi = np.full(61, 1 / 61)
    else:
        pi = np.zeros(61)
        k = {"two": [0, 1], "three": [3, 7, 19], "five": [2, 11, 29, 40, 60]}[kind]
        pi[k] = np.arange(1, len(k) + 1)
        pi /= pi.sum()
    keep = pi > 0
    for w in (0.05, 1.0, 4.0):
        _, _, R, U, V = engine.model_a_setup(2.5, 0.5, 0.3, w, w, pi)
        Q, _ = oracle.q_matrix(2.5, w, pi)
        QQ = U[0] @ np.diag(R[0]) @ V[0]
        assert np.max(np.abs((QQ - Q)[np.ix_(keep, keep)])) < 1e-12
        assert np.max(np.abs(U[0] @ V[0] - np.eye(61))) < 1e-12
        sq = np.sqrt(pi[keep])
        S = Q[np.ix_(keep, keep)] * sq[:, None] / sq[None, :]
        ev = np.linalg.eigvalsh((S + S.T) / 2)
        assert np.max(np.abs(np.sort(R[0][keep]) - ev)) < 1e-13
        if kind == "uniform":   # eigenvalues in descending order
            assert np.all(np.diff(R[0]) <= 0)


@pytest.mark.parametrize("w0,w2,same", [(0.2, 1.0, {2: 1}), (1.0, 3.0, {0: 1}),
                                        (1.0, 1.0, {0: 1, 2: 1}), (0.4, 0.4, {2: 0})])
def test_coinciding_omegas_share_one_eigen_system(w0, w2, same):
    """w2 = 1 (H0) or w0 = 1 or w0 = w2: the set takes the other's system,
    identical to computing it on its own (the same omega in a separate call)."""
    rng = np.random.default_rng(3)
    _, pi = simulate.dirichlet_f3x4(rng)
    _, _, R, U, V = engine.model_a_setup(2.2, 0.5, 0.3, w0, w2, pi)
    for s, src in same.items():
        assert np.array_equal(R[s], R[src]) and np.array_equal(U[s], U[src])
        assert np.array_equal(V[s], V[src])
    # each set equals the system of its omega computed where no sharing applies
    w = (w0, 1.0, w2)
    for s in range(3):
        a, b = (w[s], 0.123) if s == 0 else (0.123, w[s])
        _, _, Ri, Ui, Vi = engine.model_a_setup(2.2, 0.5, 0.3, a, b, pi)
        k = 0 if s == 0 else (1 if w[s] == 1.0 else 2)
        assert np.array_equal(R[s], Ri[k]) and np.array_equal(U[s], Ui[k])


def test_host_model_setup_is_thread_safe():
    """Evaluations on several host threads at once share two helper threads
    for their eigen systems (one caller at a time; the others start threads
    of their own): every thread's systems equal the ones computed alone,
    bitwise."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(11)
    _, pi = simulate.dirichlet_f3x4(rng)
    pars = [(1.5 + 0.1 * k, 0.5, 0.3, 0.05 + 0.02 * k, 2.0 + 0.3 * k) for k in range(16)]
    alone = [engine.model_a_setup(*p, pi) for p in pars]
    with ThreadPoolExecutor(8) as ex:
        for _ in range(3):
            got = list(ex.map(lambda p: engine.model_a_setup(*p, pi), pars))
            for a, b in zip(alone, got):
                for u, v in zip(a, b):
                    assert np.array_equal(u, v)


def test_noalloc_entry_refuses_genes_beyond_its_array():
    """bsa_lnl_noalloc checks the size of the gene's buffers against its
    fixed array (BSA_NOALLOC_BYTES) before any HIP call: a gene beyond it is
    refused with BSA_EINVAL and a message giving the bytes it needs."""
    rng = np.random.default_rng(4)
    tree = simulate.yule_tree(96, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, 3000, 2.0, 0.1, 3.0, 0.5, 0.3, pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), 2.0, 0.5, 0.3, 0.1, 3.0)
    with pytest.raises(engine.BsaError, match=r"needs \d+ bytes of device scratch"):
        engine.lnl_noalloc(tree, data, x)
    # and a wrong parameter count, also before any HIP call
    with pytest.raises(engine.BsaError, match="np="):
        engine.lnl_noalloc(tree, data, x[:-1])
    src = open(os.path.join(ROOT, "include", "bsa.h")).read()
    assert re.search(r"#define BSA_NOALLOC_BYTES \(\(size_t\)128 << 20\)", src)
    assert engine.NOALLOC_BYTES == 128 << 20
