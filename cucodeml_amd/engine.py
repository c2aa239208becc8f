"""Python front end (ctypes) of the C ABI in include/bsa.h.

``BranchSiteEngine`` mirrors the reference's likelihood call surface for
branch-site model A: ``lfundG(x, np)`` returns -lnL at codeml's parameter
vector x (CUDA_lfundG, treesub.c:7397), ``fhK`` holds the per-class
log f(x_h | class) (com.fhK, cuda-codeml.cu:1328) and ``gradientB`` the
finite-difference gradient ming2 uses (tools.c:6326-6350).  There is no
fallback: if libbsa.so is missing or a HIP call fails, the call raises.
"""
import ctypes as C
import os

import numpy as np

from . import codons

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libbsa.so")

BSA_FP64, BSA_FP32 = 0, 1
BSA_KERNEL_MFMA, BSA_KERNEL_VALU, BSA_KERNEL_MFMA32, BSA_KERNEL_FP32MC = 0, 1, 2, 3
BSA_PRUNE_LEVELS, BSA_PRUNE_WALK = 0, 1
PRUNE_MODES = {"levels": BSA_PRUNE_LEVELS, "walk": BSA_PRUNE_WALK}

EXPORTED = ["bsa_create", "bsa_destroy", "bsa_set_model", "bsa_num_params", "bsa_lnl",
            "bsa_lnl_grad", "bsa_lnl_grad_fd", "bsa_class_lnl", "bsa_last_params", "bsa_timing",
            "bsa_last_error", "bsa_version", "bsa_debug_mfma_probe", "bsa_work",
            "bsa_model_a_setup", "bsa_default_bounds", "bsa_fit", "bsa_debug_minimize",
            "bsa_beb", "bsa_set_graph", "bsa_batch_create", "bsa_batch_destroy",
            "bsa_batch_size", "bsa_batch_set_model", "bsa_batch_num_params",
            "bsa_batch_default_bounds", "bsa_batch_lnl", "bsa_batch_lnl_grad",
            "bsa_batch_eval_genes", "bsa_batch_fit", "bsa_batch_fit_lrt",
            "bsa_batch_class_lnl", "bsa_batch_last_params", "bsa_batch_timing", "bsa_batch_work",
            "bsa_batch_set_graph", "bsa_debug_forest_plan", "bsa_set_shapes", "bsa_tune",
            "bsa_batch_tune", "bsa_lnl_noalloc", "bsa_set_prune", "bsa_batch_set_prune"]


class BsaError(RuntimeError):
    pass


class _Problem(C.Structure):
    _fields_ = [("ns", C.c_int), ("nnode", C.c_int), ("root", C.c_int),
                ("father", C.POINTER(C.c_int)), ("label", C.POINTER(C.c_int)),
                ("ibranch", C.POINTER(C.c_int)), ("npatt", C.c_int),
                ("z", C.POINTER(C.c_ubyte)), ("ncodes", C.c_int),
                ("code_mask", C.POINTER(C.c_ubyte)), ("fpatt", C.POINTER(C.c_double)),
                ("pi", C.POINTER(C.c_double)), ("precision", C.c_int), ("kernel", C.c_int),
                ("device", C.c_int), ("tile", C.c_int)]


class _Opts(C.Structure):
    _fields_ = [("fix_omega", C.c_int), ("omega_fix", C.c_double), ("p_transformed", C.c_int)]


class _FitOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("e", C.c_double), ("max_step", C.c_double),
                ("gradient", C.c_int), ("newton", C.c_int)]


BSA_GRAD_ADJOINT, BSA_GRAD_FD = 0, 1
# why a fit stopped (BSA_STOP_*); "hend", "pgrad" and "newton" are convergence
STOP_NAMES = {0: "limit", 1: "hend", 2: "pgrad", 3: "stall", 4: "nodescent", 5: "newton"}
CONVERGED_STOPS = ("hend", "pgrad", "newton")


def fit_info(iterations, evaluations, stop):
    """Info dict of a fit: iterations, evaluations, the stop reason and
    whether it is convergence (the H_end rule, a vanishing projected
    gradient or a vanishing Newton step; a stall on a flat ridge, no
    descent step or the iteration limit are not)."""
    name = STOP_NAMES.get(int(stop), "unknown")
    return {"iterations": int(iterations), "evaluations": int(evaluations), "stop": name,
            "converged": name in CONVERGED_STOPS}
_OBJECTIVE = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.c_void_p)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise BsaError(f"{LIBPATH} not built: run `python -m cucodeml_amd.build`")
        L = C.CDLL(LIBPATH)
        dp = C.POINTER(C.c_double)
        L.bsa_create.argtypes = [C.POINTER(_Problem), C.POINTER(C.c_void_p)]
        L.bsa_destroy.argtypes = [C.c_void_p]
        L.bsa_set_model.argtypes = [C.c_void_p, C.POINTER(_Opts)]
        L.bsa_num_params.argtypes = [C.c_void_p]
        L.bsa_lnl.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.bsa_lnl_grad.argtypes = [C.c_void_p, dp, C.c_int, dp, dp]
        L.bsa_lnl_grad_fd.argtypes = [C.c_void_p, dp, C.c_int, C.c_int, C.c_double, dp, dp]
        L.bsa_class_lnl.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.bsa_last_params.argtypes = [C.c_void_p, dp]
        L.bsa_timing.argtypes = [C.c_void_p, dp, C.c_int]
        L.bsa_work.argtypes = [C.c_void_p, dp]
        L.bsa_model_a_setup.argtypes = [C.c_double] * 5 + [dp] * 6
        L.bsa_last_error.restype = C.c_char_p
        L.bsa_version.restype = C.c_char_p
        L.bsa_debug_mfma_probe.argtypes = [C.c_int, dp, dp, dp, dp]
        L.bsa_default_bounds.argtypes = [C.c_void_p, dp, dp]
        L.bsa_beb.argtypes = [C.c_void_p, dp, C.c_int, dp, dp, dp, dp, C.c_void_p]
        L.bsa_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.bsa_set_prune.argtypes = [C.c_void_p, C.c_int]
        L.bsa_batch_set_prune.argtypes = [C.c_void_p, C.c_int]
        L.bsa_set_shapes.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.bsa_tune.argtypes = [C.c_void_p, dp, C.c_int, C.c_int, dp]
        L.bsa_fit.argtypes = [C.c_void_p, dp, C.c_int, dp, dp, C.POINTER(_FitOpts), dp,
                              C.POINTER(C.c_int)]
        L.bsa_debug_minimize.argtypes = [C.c_int, _OBJECTIVE, C.c_void_p, dp, dp, dp,
                                         C.POINTER(_FitOpts), C.c_int, dp, C.POINTER(C.c_int)]
        pp, dpp = C.POINTER(C.POINTER(_Problem)), C.POINTER(dp)
        ip = C.POINTER(C.c_int)
        L.bsa_batch_create.argtypes = [pp, C.c_int, C.POINTER(C.c_void_p)]
        L.bsa_batch_destroy.argtypes = [C.c_void_p]
        L.bsa_batch_size.argtypes = [C.c_void_p]
        L.bsa_batch_set_model.argtypes = [C.c_void_p, C.c_int, C.POINTER(_Opts)]
        L.bsa_batch_num_params.argtypes = [C.c_void_p, C.c_int]
        L.bsa_batch_default_bounds.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.bsa_batch_lnl.argtypes = [C.c_void_p, dpp, dp]
        L.bsa_batch_lnl_grad.argtypes = [C.c_void_p, dpp, dp, dpp]
        L.bsa_batch_fit.argtypes = [C.c_void_p, dpp, dpp, dpp, C.POINTER(_FitOpts), dp, ip]
        L.bsa_batch_fit_lrt.argtypes = [C.c_void_p, dpp, dpp, C.POINTER(_FitOpts), dp, ip]
        L.bsa_batch_eval_genes.argtypes = [C.c_void_p, ip, C.c_int, dpp, dp, C.c_void_p]
        L.bsa_batch_class_lnl.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.bsa_batch_last_params.argtypes = [C.c_void_p, C.c_int, dp]
        L.bsa_batch_timing.argtypes = [C.c_void_p, dp, C.c_int]
        L.bsa_batch_work.argtypes = [C.c_void_p, dp]
        L.bsa_batch_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.bsa_batch_tune.argtypes = [C.c_void_p, dpp, C.c_int, dp]
        L.bsa_debug_forest_plan.argtypes = [pp, C.c_int, dp]
        L.bsa_lnl_noalloc.argtypes = [C.POINTER(_Problem), C.POINTER(_Opts), dp, C.c_int, dp,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise BsaError(f"bsa error {rc}: {lib().bsa_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def x_from_params(nbranch_t, kappa, p0, p1, w0, w2=None, transformed=True):
    """codeml parameter vector [t..., kappa, x_p0, x_p1, w0 (, w2)]."""
    if transformed:
        tot = 1.0 - p0 - p1
        a, b = np.log(p0 / tot), np.log(p1 / tot)
    else:
        a, b = p0, p1
    tail = [kappa, a, b, w0] + ([] if w2 is None else [w2])
    return np.concatenate([np.asarray(nbranch_t, dtype=np.float64), tail])


def _problem(tree, data, precision, kernel, device, tile, pi, keep):
    """bsa_problem of one gene; the arrays it points to are appended to keep"""
    def iarr(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int))

    z = np.ascontiguousarray(data.z, dtype=np.uint8)
    mask = np.ascontiguousarray(data.code_table().astype(np.uint8))
    fp = np.ascontiguousarray(data.fpatt, dtype=np.float64)
    pv = np.ascontiguousarray(data.pi if pi is None else pi, dtype=np.float64)
    keep += [z, mask, fp, pv]
    lab = np.array([max(int(v), 0) for v in tree.label])
    return _Problem(ns=tree.ns, nnode=tree.nnode, root=tree.root, father=iarr(tree.father),
                    label=iarr(lab), ibranch=iarr(tree.ibranch), npatt=data.npatt,
                    z=z.ctypes.data_as(C.POINTER(C.c_ubyte)), ncodes=data.ncodes,
                    code_mask=mask.ctypes.data_as(C.POINTER(C.c_ubyte)), fpatt=_dptr(fp),
                    pi=_dptr(pv), precision=BSA_FP32 if precision == "fp32" else BSA_FP64,
                    kernel={"valu": BSA_KERNEL_VALU, "mfma32": BSA_KERNEL_MFMA32,
                            "fp32mc": BSA_KERNEL_FP32MC}.get(kernel, BSA_KERNEL_MFMA),
                    device=device, tile=tile)


def _tune_result(o):
    return {"tile": int(o[0]), "tpb": int(o[1]), "tpc": int(o[2]), "chunks": int(o[3]),
            "ms": float(o[4]), "ms_before": float(o[5]),
            "prune": "walk" if int(o[6]) == BSA_PRUNE_WALK else "levels"}


def default_kernel(precision):
    """FP32 partials: child adjoints on the FP32 matrix cores (mfma32);
    FP64: every product on the FP64 matrix cores."""
    return "mfma32" if precision == "fp32" else "mfma"


class BranchSiteEngine:
    def __init__(self, tree, data, precision="fp64", kernel=None, device=0,
                 fix_omega=False, omega_fix=1.0, p_transformed=True, pi=None, tile=0):
        kernel = kernel or default_kernel(precision)
        self.kernel = kernel
        self.tree, self.data = tree, data
        self.precision, self.device = precision, device
        self._keep = []
        prob = _problem(tree, data, precision, kernel, device, tile, pi, self._keep)
        h = C.c_void_p()
        _check(lib().bsa_create(C.byref(prob), C.byref(h)))
        self._h = h
        self.set_model(fix_omega, omega_fix, p_transformed)

    def set_model(self, fix_omega=False, omega_fix=1.0, p_transformed=True):
        o = _Opts(int(bool(fix_omega)), float(omega_fix), int(bool(p_transformed)))
        _check(lib().bsa_set_model(self._h, C.byref(o)))
        self.np = lib().bsa_num_params(self._h)
        self.fix_omega, self.omega_fix = bool(fix_omega), float(omega_fix)
        self.p_transformed = bool(p_transformed)

    def use_graph(self, enable=True):
        """Replay each evaluation's device work as a captured HIP graph."""
        _check(lib().bsa_set_graph(self._h, int(bool(enable))))

    def set_shapes(self, tile=64, tpb=0, tpc=0):
        """Launch shapes (bsa_set_shapes)."""
        _check(lib().bsa_set_shapes(self._h, tile, tpb, tpc))

    def set_prune(self, mode="levels"):
        """Pruning by one launch per level ("levels") or by the persistent
        walk ("walk"), bsa_set_prune."""
        _check(lib().bsa_set_prune(self._h, PRUNE_MODES[mode]))

    def tune(self, x, reps=3):
        """Autotune the launch shapes at x (bsa_tune); returns the choice."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        o = np.zeros(7)
        _check(lib().bsa_tune(self._h, _dptr(x), len(x), reps, _dptr(o)))
        return _tune_result(o)

    def close(self):
        if getattr(self, "_h", None):
            lib().bsa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lnl(self, x):
        """-lnL at codeml variables x (as lfundG)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = C.c_double()
        _check(lib().bsa_lnl(self._h, _dptr(x), len(x), C.byref(out)))
        return out.value

    lfundG = lnl

    def lnl_grad(self, x):
        """(-lnL, d(-lnL)/dx) by the adjoint sweep (ming2's dfun)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(len(x))
        out = C.c_double()
        _check(lib().bsa_lnl_grad(self._h, _dptr(x), len(x), C.byref(out), _dptr(g)))
        return out.value, g

    dfun = lnl_grad

    def gradientB(self, x, central=False, small=1e-6):
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(len(x))
        out = C.c_double()
        _check(lib().bsa_lnl_grad_fd(self._h, _dptr(x), len(x), int(central), small,
                                     C.byref(out), _dptr(g)))
        return out.value, g

    def default_bounds(self):
        """codeml's bounds (lo, hi) for this engine's x layout (SetxBound)."""
        lo, hi = np.zeros(self.np), np.zeros(self.np)
        _check(lib().bsa_default_bounds(self._h, _dptr(lo), _dptr(hi)))
        return lo, hi

    def fit(self, x0, lo=None, hi=None, e=1e-8, max_iter=2000, max_step=5.0,
            gradient="adjoint", newton=0):
        """Bounded BFGS maximum-likelihood fit from x0 (replaces ming2 at
        codeml.c:688).  Returns (x, -lnL, fit_info dict)."""
        x = np.array(x0, dtype=np.float64)
        if len(x) != self.np:
            raise BsaError(f"x has {len(x)} values, the engine expects {self.np}")
        dlo, dhi = self.default_bounds()
        lo = dlo if lo is None else np.ascontiguousarray(lo, dtype=np.float64)
        hi = dhi if hi is None else np.ascontiguousarray(hi, dtype=np.float64)
        o = _FitOpts(int(max_iter), float(e), float(max_step),
                     BSA_GRAD_FD if gradient == "fd" else BSA_GRAD_ADJOINT, int(newton))
        f = C.c_double()
        info = (C.c_int * 3)()
        _check(lib().bsa_fit(self._h, _dptr(x), len(x), _dptr(lo), _dptr(hi), C.byref(o),
                             C.byref(f), info))
        return x, f.value, fit_info(info[0], info[1], info[2])

    def beb(self, x, sets=False):
        """Bayes empirical Bayes at x (bsa_beb): dict with the class
        posteriors per pattern "post" (4, npatt), "log_fX", the posteriors on
        the w0 and w2 grids and on the p0-p1 triangles, and with sets=True
        the 121-set table "lnf_sets" (121, npatt)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = self.data.npatt
        post, ml, pw, p01 = np.zeros(4 * n), np.zeros(3), np.zeros(20), np.zeros(100)
        tab = np.zeros(121 * n) if sets else None
        _check(lib().bsa_beb(self._h, _dptr(x), len(x), _dptr(post), _dptr(ml), _dptr(pw),
                             _dptr(p01), None if tab is None else tab.ctypes.data))
        out = {"post": post.reshape(4, n), "log_fX": float(ml[0]), "S1": float(ml[1]),
               "log_grid_sum": float(ml[2]), "post_w0": pw[:10], "post_w2": pw[10:],
               "post_p0p1": p01}
        if sets:
            out["lnf_sets"] = tab.reshape(121, n)
        return out

    def neb(self, x):
        """Naive empirical Bayes class posteriors per pattern (4, npatt) at x:
        p_k f(x_h | k) / f(x_h) (lfunNSsites_rate, codeml.c:5079-5129)."""
        self.lnl(x)
        fh, site = self.fhK()
        fk = self.last_params()[7:11]
        return fk[:, None] * np.exp(fh - site[None, :])

    def fhK(self):
        """(per-class log f (4, npatt), per-pattern mixture log f (npatt,))"""
        f = np.zeros(4 * self.data.npatt)
        s = np.zeros(self.data.npatt)
        _check(lib().bsa_class_lnl(self._h, f.ctypes.data, s.ctypes.data))
        return f.reshape(4, -1), s

    def last_params(self):
        o = np.zeros(11)
        _check(lib().bsa_last_params(self._h, _dptr(o)))
        return o

    def work(self):
        o = np.zeros(10)
        _check(lib().bsa_work(self._h, _dptr(o)))
        return o

    def timing(self, reset=False):
        """dict of last / summed device times (ms) and counts"""
        o = np.zeros(17)
        _check(lib().bsa_timing(self._h, _dptr(o), int(reset)))
        keys = ["pmat", "prune", "root", "adjoint", "branch", "finish", "wall"]
        d = {f"last_{k}": float(o[i]) for i, k in enumerate(keys)}
        d.update({f"sum_{k}": float(o[7 + i]) for i, k in enumerate(keys)})
        d["n_eval"], d["n_launch"], d["n_adj_launch"] = float(o[14]), float(o[15]), float(o[16])
        return d


def _ptrs(arrays):
    """C array of double pointers to the given float64 arrays"""
    return (C.POINTER(C.c_double) * len(arrays))(*[_dptr(a) for a in arrays])


class BranchSiteBatch:
    """Many genes evaluated together (bsa_batch_*): each gene has its own
    tree, alignment and model options; one evaluation of the batch runs one
    launch per tree level of the deepest tree for all of them.  Patterns are
    padded to the largest gene's count, so genes of similar size batch best.
    ``genes`` is a list of (tree, data) pairs."""

    def __init__(self, genes, precision="fp64", kernel=None, device=0, tile=0,
                 fix_omega=False, omega_fix=1.0, p_transformed=True):
        kernel = kernel or default_kernel(precision)
        self.kernel = kernel
        self.genes = list(genes)
        self.precision, self.device = precision, device
        self._keep = []
        probs = [_problem(t, d, precision, kernel, device, tile, None, self._keep)
                 for t, d in self.genes]
        self._probs = probs
        arr = (C.POINTER(_Problem) * len(probs))(*[C.pointer(p) for p in probs])
        h = C.c_void_p()
        _check(lib().bsa_batch_create(arr, len(probs), C.byref(h)))
        self._h = h
        self.np = [0] * len(self.genes)
        self.fix_omega = [False] * len(self.genes)
        self.omega_fix = [1.0] * len(self.genes)
        self.set_model(-1, fix_omega, omega_fix, p_transformed)

    def __len__(self):
        return len(self.genes)

    def set_model(self, gene=-1, fix_omega=False, omega_fix=1.0, p_transformed=True):
        """model options of one gene, or of every gene with gene < 0"""
        o = _Opts(int(bool(fix_omega)), float(omega_fix), int(bool(p_transformed)))
        _check(lib().bsa_batch_set_model(self._h, int(gene), C.byref(o)))
        for i in (range(len(self.genes)) if gene < 0 else [gene]):
            self.np[i] = lib().bsa_batch_num_params(self._h, i)
            self.fix_omega[i], self.omega_fix[i] = bool(fix_omega), float(omega_fix)

    def use_graph(self, enable=True):
        _check(lib().bsa_batch_set_graph(self._h, int(bool(enable))))

    def set_prune(self, mode="levels"):
        """Pruning by level launches or by the persistent walk
        (bsa_batch_set_prune)."""
        _check(lib().bsa_batch_set_prune(self._h, PRUNE_MODES[mode]))

    def close(self):
        if getattr(self, "_h", None):
            lib().bsa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _xs(self, xs):
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        if len(xs) != len(self.genes):
            raise BsaError(f"{len(xs)} parameter vectors for {len(self.genes)} genes")
        for i, x in enumerate(xs):
            if len(x) != self.np[i]:
                raise BsaError(f"gene {i}: x has {len(x)} values, expected {self.np[i]}")
        return xs

    def lnl(self, xs):
        """-lnL of every gene at xs[gene]"""
        xs = self._xs(xs)
        out = np.zeros(len(xs))
        _check(lib().bsa_batch_lnl(self._h, _ptrs(xs), _dptr(out)))
        return out

    def lnl_grad(self, xs):
        """(-lnL per gene, list of d(-lnL)/dx per gene) by the adjoint sweep"""
        xs = self._xs(xs)
        out = np.zeros(len(xs))
        gs = [np.zeros(len(x)) for x in xs]
        _check(lib().bsa_batch_lnl_grad(self._h, _ptrs(xs), _dptr(out), _ptrs(gs)))
        return out, gs

    def tune(self, xs, reps=3):
        """Autotune the batch's launch shapes at xs (bsa_batch_tune)."""
        xs = self._xs(xs)
        o = np.zeros(7)
        _check(lib().bsa_batch_tune(self._h, _ptrs(xs), reps, _dptr(o)))
        return _tune_result(o)

    def eval_genes(self, genes, xs, grad=True):
        """-lnL (and gradients) of the listed genes only (ascending indices)
        at xs[k] for genes[k]; the other genes keep their last results."""
        genes = np.ascontiguousarray(genes, dtype=np.int32)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        out = np.zeros(len(genes))
        gs = [np.zeros(len(x)) for x in xs] if grad else None
        _check(lib().bsa_batch_eval_genes(self._h, genes.ctypes.data_as(C.POINTER(C.c_int)),
                                          len(genes), _ptrs(xs), _dptr(out),
                                          C.cast(_ptrs(gs), C.c_void_p) if grad else None))
        return (out, gs) if grad else out

    def default_bounds(self, gene):
        lo, hi = np.zeros(self.np[gene]), np.zeros(self.np[gene])
        _check(lib().bsa_batch_default_bounds(self._h, gene, _dptr(lo), _dptr(hi)))
        return lo, hi

    def fit(self, x0s, e=1e-8, max_iter=2000, max_step=5.0, lo=None, hi=None, newton=0):
        """ML fits of every gene from x0s[gene] (bounded BFGS, adjoint
        gradient), one optimiser thread per gene, served in rounds of the
        genes whose points are ready.  lo / hi: per-gene bound vectors (None,
        or None entries: codeml's bounds); max_iter < 0 and newton > 0:
        Newton iterations only (bsa_fit_opts).  Returns (list of x, -lnL
        per gene, list of info dicts)."""
        xs = [np.array(x, dtype=np.float64) for x in self._xs(x0s)]
        o = _FitOpts(int(max_iter), float(e), float(max_step), BSA_GRAD_ADJOINT, int(newton))
        out = np.zeros(len(xs))
        info = np.zeros(3 * len(xs), dtype=np.int32)

        def bound_ptrs(b):
            if b is None:
                return None
            arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in b]
            keep.append(arrs)
            dp = C.POINTER(C.c_double)
            return (dp * len(arrs))(*[dp() if a is None else _dptr(a) for a in arrs])
        keep = []
        _check(lib().bsa_batch_fit(self._h, _ptrs(xs), bound_ptrs(lo), bound_ptrs(hi),
                                   C.byref(o), _dptr(out),
                                   info.ctypes.data_as(C.POINTER(C.c_int))))
        infos = [fit_info(info[3 * i], info[3 * i + 1], info[3 * i + 2])
                 for i in range(len(xs))]
        return xs, out, infos

    def fit_lrt(self, x1s, e=1e-8, max_iter=2000, max_step=5.0):
        """H1, then H0 from the H1 estimates, then (where H0 lies above H1)
        H1 again from the H0 estimates, for every gene, each gene's chain in
        its own optimiser thread (bsa_batch_fit_lrt).  Every gene's model is
        left at H1.  Returns (list of H1 x, list of H0 x, -lnL of H1 and of
        H0 per gene (n, 2), list of info dicts)."""
        self.set_model(-1, False)
        x1 = [np.array(x, dtype=np.float64) for x in self._xs(x1s)]
        x0 = [np.zeros(len(x) - 1) for x in x1]
        o = _FitOpts(int(max_iter), float(e), float(max_step), BSA_GRAD_ADJOINT, 0)
        out = np.zeros(2 * len(x1))
        info = np.zeros(7 * len(x1), dtype=np.int32)
        _check(lib().bsa_batch_fit_lrt(self._h, _ptrs(x1), _ptrs(x0), C.byref(o), _dptr(out),
                                       info.ctypes.data_as(C.POINTER(C.c_int))))
        infos = []
        for i in range(len(x1)):
            o = info[7 * i:7 * i + 7]
            d = {"refit": int(o[6])}
            for hyp, k in (("h1", 0), ("h0", 3)):
                for key, v in fit_info(o[k], o[k + 1], o[k + 2]).items():
                    d[f"{hyp}_{key}"] = v
            infos.append(d)
        return x1, x0, out.reshape(-1, 2), infos

    def fhK(self, gene):
        n = self.genes[gene][1].npatt
        f, s = np.zeros(4 * n), np.zeros(n)
        _check(lib().bsa_batch_class_lnl(self._h, gene, f.ctypes.data, s.ctypes.data))
        return f.reshape(4, -1), s

    def last_params(self, gene):
        o = np.zeros(11)
        _check(lib().bsa_batch_last_params(self._h, gene, _dptr(o)))
        return o

    def work(self):
        o = np.zeros(10)
        _check(lib().bsa_batch_work(self._h, _dptr(o)))
        return o

    def timing(self, reset=False):
        o = np.zeros(17)
        _check(lib().bsa_batch_timing(self._h, _dptr(o), int(reset)))
        keys = ["pmat", "prune", "root", "adjoint", "branch", "finish", "wall"]
        d = {f"last_{k}": float(o[i]) for i, k in enumerate(keys)}
        d.update({f"sum_{k}": float(o[7 + i]) for i, k in enumerate(keys)})
        d["n_eval"], d["n_launch"], d["n_adj_launch"] = float(o[14]), float(o[15]), float(o[16])
        return d


# bytes of the device array bsa_lnl_noalloc works in (BSA_NOALLOC_BYTES)
NOALLOC_BYTES = 128 << 20


def lnl_noalloc(tree, data, x, precision="fp64", kernel=None, device=0, fix_omega=False,
                omega_fix=1.0, p_transformed=True, stream=None, tile=0):
    """-lnL, per-class log f (4, npatt) and per-site log f of one gene by
    bsa_lnl_noalloc: no engine, no device allocation (every device buffer in
    the library's fixed array on the device).  stream: a hipStream_t value
    (int) or None for the null stream."""
    keep = []
    prob = _problem(tree, data, precision, kernel or default_kernel(precision), device, tile,
                    None, keep)
    o = _Opts(int(bool(fix_omega)), float(omega_fix), int(bool(p_transformed)))
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = C.c_double()
    fh, site = np.zeros(4 * data.npatt), np.zeros(data.npatt)
    _check(lib().bsa_lnl_noalloc(C.byref(prob), C.byref(o), _dptr(x), len(x), C.byref(out),
                                 fh.ctypes.data, site.ctypes.data, stream))
    return out.value, fh.reshape(4, -1), site


def forest_plan(genes, precision="fp64"):
    """Host-only plan of a batch of (tree, data) genes, every schedule index
    checked (bsa_debug_forest_plan): dict of its sizes."""
    keep = []
    probs = [_problem(t, d, precision, "mfma", 0, 0, None, keep) for t, d in genes]
    arr = (C.POINTER(_Problem) * len(probs))(*[C.pointer(p) for p in probs])
    o = np.zeros(8)
    _check(lib().bsa_debug_forest_plan(arr, len(probs), _dptr(o)))
    keys = ["genes", "npad", "nslot", "levels", "prune_jobs", "adj_jobs", "eig_jobs", "nchunk"]
    return {k: int(v) for k, v in zip(keys, o)}


def minimize(fun, x0, lo, hi, e=1e-8, max_iter=2000, max_step=5.0, newton=0, block=0):
    """The engine's bounded BFGS on a Python objective fun(x) -> (f, grad)
    (diagnostic; runs on the host).  block: the last `block` variables,
    whose Hessian block Newton steps on more than 160 free variables form
    exactly (bsa_debug_minimize).  Returns (x, f, info)."""
    n = len(x0)

    def cb(xp, fp, gp, _ctx):
        try:
            xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
            f, g = fun(xv)
            if not np.isfinite(f):
                return 1
            fp[0] = float(f)
            for i in range(n):
                gp[i] = float(g[i])
            return 0
        except Exception:
            return 1

    cfun = _OBJECTIVE(cb)
    x = np.array(x0, dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    o = _FitOpts(int(max_iter), float(e), float(max_step), BSA_GRAD_ADJOINT, int(newton))
    f = C.c_double()
    info = (C.c_int * 3)()
    _check(lib().bsa_debug_minimize(n, cfun, None, _dptr(x), _dptr(lo), _dptr(hi), C.byref(o),
                                    int(block), C.byref(f), info))
    return x, f.value, fit_info(info[0], info[1], info[2])


def model_a_setup(kappa, p0, p1, w0, w2, pi):
    """Host-side model set-up: (freqK[4], Qfactor[2], Root[3,61], U[3,61,61], V[3,61,61])."""
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    fk, qf = np.zeros(4), np.zeros(2)
    R, U, V = np.zeros(3 * 61), np.zeros(3 * 61 * 61), np.zeros(3 * 61 * 61)
    _check(lib().bsa_model_a_setup(kappa, p0, p1, w0, w2, _dptr(pi), _dptr(fk), _dptr(qf),
                                   _dptr(R), _dptr(U), _dptr(V)))
    return fk, qf, R.reshape(3, 61), U.reshape(3, 61, 61), V.reshape(3, 61, 61)


def mfma_probe(A, B, device=0):
    A = np.ascontiguousarray(A, dtype=np.float64).ravel()
    B = np.ascontiguousarray(B, dtype=np.float64).ravel()
    d64, d32 = np.zeros(256), np.zeros(256)
    _check(lib().bsa_debug_mfma_probe(device, _dptr(A), _dptr(B), _dptr(d64), _dptr(d32)))
    return d64.reshape(64, 4), d32.reshape(64, 4)


__all__ = ["BranchSiteEngine", "BranchSiteBatch", "BsaError", "x_from_params", "mfma_probe", "lib", "codons"]
