"""ORACLE — test infrastructure only (the checker, never the product).

ctypes front end of bsa_oracle.c, the plain-C FP64 restatement of the
reference's branch-site model A likelihood (see the header of bsa_oracle.c
for the reference file:line each function follows).  Only tests/, the smoke
check and bench.py's CPU-baseline leg may import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_CODEML = os.path.join(HERE, "_ref", "codeml")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_lib = None


def build():
    src = os.path.join(HERE, "bsa_oracle.c")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-o", LIB, src, "-lm"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_q.argtypes = [C.c_double, C.c_double, _dp, _dp, C.POINTER(C.c_double)]
        L.orc_eigen.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_pmat.argtypes = [_dp, _dp, _dp, C.c_double, _dp]
        L.orc_model.argtypes = [C.c_double] * 5 + [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_scale_marks.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, C.c_char_p]
        L.orc_scale_marks.restype = C.c_int
        L.orc_lnl.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, C.c_int,
                              _up, C.c_int, _dp, _dp, _dp] + [C.c_double] * 5 + \
                             [C.c_int, C.c_void_p, C.c_void_p]
        L.orc_lnl.restype = C.c_double
        L.orc_lnl_qf.argtypes = L.orc_lnl.argtypes[:20] + [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_lnl_qf.restype = C.c_double
        _lib = L
    return _lib


def q_matrix(kappa, omega, pi):
    Q = np.zeros(61 * 61)
    mr = C.c_double()
    lib().orc_q(kappa, omega, np.ascontiguousarray(pi, dtype=np.float64), Q, C.byref(mr))
    return Q.reshape(61, 61), mr.value


def eigen(Q, pi):
    R, U, V = np.zeros(61), np.zeros(61 * 61), np.zeros(61 * 61)
    lib().orc_eigen(np.ascontiguousarray(Q, dtype=np.float64).ravel(),
                    np.ascontiguousarray(pi, dtype=np.float64), R, U, V)
    return R, U.reshape(61, 61), V.reshape(61, 61)


def pmat(U, V, Root, t):
    P = np.zeros(61 * 61)
    lib().orc_pmat(np.ascontiguousarray(U).ravel(), np.ascontiguousarray(V).ravel(),
                   np.ascontiguousarray(Root), float(t), P)
    return P.reshape(61, 61)


def model(kappa, p0, p1, w0, w2, pi):
    fk, qf = np.zeros(4), np.zeros(2)
    R3, U3, V3 = np.zeros(3 * 61), np.zeros(3 * 61 * 61), np.zeros(3 * 61 * 61)
    lib().orc_model(kappa, p0, p1, w0, w2, np.ascontiguousarray(pi, dtype=np.float64),
                    fk, qf, R3, U3, V3)
    return fk, qf, R3.reshape(3, 61), U3.reshape(3, 61, 61), V3.reshape(3, 61, 61)


def _tree_arrays(tree):
    maxson = max(len(s) for s in tree.sons)
    nson = np.array([len(s) for s in tree.sons], dtype=np.int32)
    sons = np.zeros((tree.nnode, maxson), dtype=np.int32)
    for v, s in enumerate(tree.sons):
        sons[v, :len(s)] = s
    lab = np.array([max(int(l), 0) for l in tree.label], dtype=np.int32)
    return maxson, nson, np.ascontiguousarray(sons.ravel()), lab


def scale_marks(tree):
    maxson, nson, sons, _ = _tree_arrays(tree)
    buf = C.create_string_buffer(tree.nnode)
    n = lib().orc_scale_marks(tree.nnode, tree.root, nson, sons, maxson, buf)
    return n, np.frombuffer(buf.raw, dtype=np.uint8).copy()


def lnl(tree, data, kappa, p0, p1, w0, w2, use_scaling=True, pi=None, qfactor=None):
    """lnL, per-class log f (4, npatt) and per-pattern mixture log f.
    ``qfactor`` (background, foreground) replaces the model's Qfactors."""
    maxson, nson, sons, lab = _tree_arrays(tree)
    pi = data.pi if pi is None else pi
    fhK = np.zeros(4 * data.npatt)
    site = np.zeros(data.npatt)
    z = np.ascontiguousarray(data.z, dtype=np.uint8)
    mask = np.ascontiguousarray(data.code_table())
    qf = None if qfactor is None else np.ascontiguousarray(qfactor, dtype=np.float64)
    v = lib().orc_lnl_qf(tree.ns, tree.nnode, tree.root, maxson, nson, sons, lab,
                         np.ascontiguousarray(tree.branch, dtype=np.float64), data.npatt, z,
                         mask.shape[0], mask.ravel(), np.ascontiguousarray(data.fpatt),
                         np.ascontiguousarray(pi, dtype=np.float64),
                         kappa, p0, p1, w0, w2, int(use_scaling),
                         None if qf is None else qf.ctypes.data, fhK.ctypes.data,
                         site.ctypes.data)
    return v, fhK.reshape(4, data.npatt), site


# ---------------------------------------------------------------- site posteriors

BEB_N1D = 10


def neb(tree, data, kappa, p0, p1, w0, w2):
    """Naive empirical Bayes class posteriors per pattern (4, npatt) at the
    given parameters (lfunNSsites_rate, codeml.c:5079-5129): p_k f_k / f."""
    _, fh, site = lnl(tree, data, kappa, p0, p1, w0, w2)
    t = p0 + p1
    fk = np.array([p0, p1, (1 - t) * p0 / t, (1 - t) * p1 / t])
    return fk[:, None] * np.exp(fh - site[None, :])


def beb_sets(tree, data, kappa, p0, p1, w0, w2):
    """log f(x_h | w set) for the 121 w sets of model A's BEB, in the
    reference's order (get_grid_para_like_ACD, codeml.c:6419-6607): branch
    lengths, kappa and the Qfactors of the given parameters held fixed."""
    n1d = BEB_N1D
    fk, qf, *_ = model(kappa, p0, p1, w0, w2, data.pi)
    w0g = (np.arange(n1d) + 0.5) / n1d
    w2g = 1 + (np.arange(n1d) + 0.5) * (11 - 1) / n1d
    F = np.zeros((n1d + 1 + n1d * n1d + n1d, data.npatt))
    for i in range(n1d):
        for j in range(n1d):
            _, fh, _ = lnl(tree, data, kappa, p0, p1, w0g[i], w2g[j], qfactor=qf)
            F[n1d + 1 + i * n1d + j] = fh[2]
            if j == 0:
                F[i] = fh[0]
            if i == 0:
                F[n1d + 1 + n1d * n1d + j] = fh[3]
            if i == 0 and j == 0:
                F[n1d] = fh[1]
    return F


def beb_grid():
    """(pclassM[ngrid, 4], iw[ngrid, 4]) of model A's 10^4-point grid
    (get_pclassM_iw_ACD codeml.c:6609-6615, GetIndexTernary tools.c:4441-4461)."""
    n1d = BEB_N1D
    g = np.arange(n1d ** 4)
    ip = np.stack([g // 1000, (g // 100) % 10, (g // 10) % 10, g % 10], axis=1)
    tri = ip[:, 0] * n1d + ip[:, 1]
    ix = np.floor(np.sqrt(tri)).astype(int)
    iy = tri - ix * ix
    q0 = (1 + (iy // 2) * 3 + iy % 2) / (3.0 * n1d)
    q1 = (1 + (n1d - 1 - ix) * 3 + iy % 2) / (3.0 * n1d)
    q2 = 1 - q0 - q1
    pc = np.stack([q0, q1, q2 * q0 / (1 - q2), q2 * q1 / (1 - q2)], axis=1)
    iw = np.stack([ip[:, 2], np.full_like(g, n1d), n1d + 1 + ip[:, 2] * n1d + ip[:, 3],
                   n1d + 1 + n1d * n1d + ip[:, 3]], axis=1)
    return pc, iw, ip


def beb_integrate(F, fpatt):
    """The grid integration of lfunNSsites_ACD (codeml.c:6699-6769) on a
    121-set table of log f: returns (post (4, npatt), log f(X) as printed,
    posterior on the w0 and w2 grids (2, 10), on the p0-p1 triangles (100,))."""
    n1d, dim = BEB_N1D, 4
    mx = F.max(axis=0)
    S1 = float(np.sum(fpatt * mx))
    E = np.exp(F - mx[None, :])
    pc, iw, ip = beb_grid()
    ngrid = pc.shape[0]
    lnfx = np.zeros(ngrid)
    for g in range(ngrid):
        fh = (pc[g][:, None] * E[iw[g]]).sum(axis=0)
        ok = fh >= 1e-300
        lnfx[g] = np.sum(np.log(fh[ok]) * fpatt[ok])
    fX, S2 = 1.0, -1e300
    pp = np.zeros((dim, n1d))
    p01 = np.zeros(n1d * n1d)
    for g in range(ngrid):
        t = lnfx[g] - S2
        if t > 10:
            t = np.exp(-t) if t < 200 else 0.0
            fX = fX * t + 1
            pp *= t
            p01 *= t
            pp[np.arange(dim), ip[g]] += 1
            p01[ip[g, 0] * n1d + ip[g, 1]] += 1
            S2 = lnfx[g]
        elif t > -200:
            t = np.exp(t)
            fX += t
            pp[np.arange(dim), ip[g]] += t
            p01[ip[g, 0] * n1d + ip[g, 1]] += t
    pp /= fX
    p01 /= fX
    lfX = np.log(fX) + S2
    post = np.zeros((4, F.shape[1]))
    with np.errstate(divide="ignore", invalid="ignore"):
        for g in range(ngrid):
            fk = pc[g][:, None] * E[iw[g]]
            fh = fk.sum(axis=0)
            t = np.log(fk / fh) + lnfx[g] - lfX
            post += np.where(t > -300, np.exp(t), 0.0)
    return post, lfX + S1 - dim * np.log(n1d), pp[2:], p01


# ---------------------------------------------------------------- standard errors

def standard_errors(tree, data, x, fix_omega, omega_fix=1.0, small_diff=1e-6):
    """SEs of codeml's parameters at x (branch lengths, kappa, p0, p1, w0
    (, w2), p's untransformed) from the Hessian approximation of Seo et al.
    2004 as HessianSKT2004 computes it (treesub.c:6979-7048): per-site central
    differences with step 2 Small_Diff (|x_i| + 1), H = -sum_h fpatt_h
    d_i d_j, SE_i = sqrt([(-H)^-1]_ii) (-1 where not positive;
    codeml.c:779-784)."""
    nb = tree.nbranch
    x = np.asarray(x, dtype=np.float64)

    def site(xv):
        tree.set_branch_vector(xv[:nb])
        w2 = omega_fix if fix_omega else xv[nb + 4]
        _, _, s = lnl(tree, data, xv[nb], xv[nb + 1], xv[nb + 2], xv[nb + 3], w2)
        return s

    eh0 = small_diff * 2
    D = np.zeros((len(x), data.npatt))
    for i in range(len(x)):
        eh = eh0 * (abs(x[i]) + 1)
        xm, xp = x.copy(), x.copy()
        xm[i] -= eh
        xp[i] += eh
        D[i] = (site(xp) - site(xm)) / (2 * eh)
    tree.set_branch_vector(x[:nb])
    H = -(D * data.fpatt[None, :]) @ D.T
    V = np.linalg.inv(-H)
    d = np.diag(V)
    return np.where(d > 0, np.sqrt(np.abs(d)), -1.0)
