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


def lnl(tree, data, kappa, p0, p1, w0, w2, use_scaling=True, pi=None):
    """lnL, per-class log f (4, npatt) and per-pattern mixture log f."""
    maxson, nson, sons, lab = _tree_arrays(tree)
    pi = data.pi if pi is None else pi
    fhK = np.zeros(4 * data.npatt)
    site = np.zeros(data.npatt)
    z = np.ascontiguousarray(data.z, dtype=np.uint8)
    mask = np.ascontiguousarray(data.code_table())
    v = lib().orc_lnl(tree.ns, tree.nnode, tree.root, maxson, nson, sons, lab,
                      np.ascontiguousarray(tree.branch, dtype=np.float64), data.npatt, z,
                      mask.shape[0], mask.ravel(), np.ascontiguousarray(data.fpatt),
                      np.ascontiguousarray(pi, dtype=np.float64),
                      kappa, p0, p1, w0, w2, int(use_scaling),
                      fhK.ctypes.data, site.ctypes.data)
    return v, fhK.reshape(4, data.npatt), site
