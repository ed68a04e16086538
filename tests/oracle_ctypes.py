"""ctypes view of the CPU oracle (oracle/liboracle.so) for the test-suite.

The oracle is the checker: tests, the smoke check and the benchmark's CPU
baseline load it; the product never does.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class OrMatrix(C.Structure):
    _fields_ = [("n", C.c_int), ("ncomp", C.c_int), ("rowptr", i32p), ("col", i32p),
                ("val", f64p), ("diag", i32p)]


class OrTransfer(C.Structure):
    _fields_ = [("ncoarse", C.c_int), ("nfine", C.c_int),
                ("crow", i32p), ("ccol", i32p), ("cval", f64p),
                ("frow", i32p), ("fcol", i32p), ("fval", f64p)]


class OrHN(C.Structure):
    _fields_ = [("kind", C.c_int), ("ne", C.c_int), ("nf", C.c_int),
                ("edges", i32p), ("faces", i32p)]


class OrCGInfo(C.Structure):
    _fields_ = [("tol", C.c_double), ("globaltol", C.c_double), ("breaktol", C.c_double),
                ("miniter", C.c_int), ("maxiter", C.c_int),
                ("status", C.c_int), ("iteration", C.c_int), ("totaliter", C.c_int),
                ("firstresidual", C.c_double), ("residual", C.c_double),
                ("aimedresidual", C.c_double), ("previousresidual", C.c_double),
                ("correction", C.c_double), ("previouscorrection", C.c_double),
                ("rate", C.c_double), ("lastrate", C.c_double)]


class OrNLInfo(C.Structure):
    _fields_ = [("tol", C.c_double), ("globaltol", C.c_double), ("breaktol", C.c_double),
                ("rho", C.c_double), ("linrho", C.c_double), ("maxresincrease", C.c_double),
                ("miniter", C.c_int), ("maxiter", C.c_int), ("maxrelax", C.c_int),
                ("status", C.c_int), ("iteration", C.c_int), ("laststepbad", C.c_int),
                ("relax", C.c_int), ("totaliter", C.c_int),
                ("firstresidual", C.c_double), ("residual", C.c_double),
                ("aimedresidual", C.c_double), ("previousresidual", C.c_double),
                ("correction", C.c_double), ("previouscorrection", C.c_double),
                ("rate", C.c_double), ("lastrate", C.c_double)]


class OrIlu(C.Structure):
    _fields_ = [("n", C.c_int), ("ncomp", C.c_int), ("p", i32p), ("q", i32p),
                ("rowptr", i32p), ("col", i32p), ("diag", i32p), ("val", f64p), ("y", f64p)]


class OrLevel(C.Structure):
    _fields_ = [("A", OrMatrix), ("T", C.POINTER(OrTransfer)), ("hn", OrHN),
                ("ndir", C.c_int), ("dir", i32p), ("ndircomp", C.c_int), ("dircomp", i32p),
                ("direct", C.c_int), ("lu", f64p), ("piv", i32p), ("vanka", C.c_void_p),
                ("ilu", C.c_void_p)]


class OrMG(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("L", C.POINTER(OrLevel)), ("mgtype", C.c_char),
                ("iterpre", C.c_int), ("iterpost", C.c_int), ("iterexact", C.c_int),
                ("omega", C.c_double), ("mgomega", C.c_double), ("smoother", C.c_int),
                ("prec_maxiter", C.c_int), ("prec_tol", C.c_double),
                ("prec_globaltol", C.c_double), ("mgsteps", C.c_int64),
                ("u", C.POINTER(f64p)), ("b", C.POINTER(f64p)), ("v", C.POINTER(f64p)),
                ("plant", C.c_int)]


class OrSolveReport(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("gmres_iterations", C.c_int),
                ("mgsteps", C.c_int64), ("status", C.c_int),
                ("first_residual", C.c_double), ("last_residual", C.c_double),
                ("damping_steps", C.c_int), ("damping_exploded", C.c_int)]


# planted-mistake switches of or_mg.plant (gascoigne_oracle.h)
PLANT_NO_HN_DISTRIBUTE, PLANT_NO_HN_AVERAGE, PLANT_NO_HN_ZERO_FINE = 1, 2, 4

CG_STATUS = {0: "waiting", 1: "running", 2: "converged", 3: "too many iterations",
             4: "exploded", 5: "diverging"}
NL_STATUS = {0: "waiting", 1: "running", 2: "converged", 3: "diverged",
             4: "slow_convergence", 5: "stagnation", 6: "exploded"}


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_matrix_init": (C.c_int, [C.POINTER(OrMatrix), C.c_int, C.c_int, i32p, i32p, f64p]),
        "or_matrix_free": (None, [C.POINTER(OrMatrix)]),
        "or_vmult": (None, [C.POINTER(OrMatrix), f64p, f64p, C.c_double]),
        "or_vmult_transpose": (None, [C.POINTER(OrMatrix), f64p, f64p, C.c_double]),
        "or_residual": (None, [C.POINTER(OrMatrix), f64p, f64p, f64p]),
        "or_jacobi": (None, [C.POINTER(OrMatrix), f64p]),
        "or_block_gauss_jordan": (None, [C.c_int, f64p]),
        "or_norm": (C.c_double, [f64p, C.c_int64]),
        "or_norm_l8": (C.c_double, [f64p, C.c_int64]),
        "or_dot": (C.c_double, [f64p, f64p, C.c_int64]),
        "or_transfer_create": (C.POINTER(OrTransfer), [C.c_int, i32p, C.c_int, i32p, C.c_int,
                                                       i32p, C.c_int, i32p]),
        "or_transfer_free": (None, [C.POINTER(OrTransfer)]),
        "or_transfer_nnz": (C.c_int64, [C.POINTER(OrTransfer)]),
        "or_transfer_export": (None, [C.POINTER(OrTransfer), i32p, i32p, f64p, i32p, i32p, f64p]),
        "or_restrict_zero": (None, [C.POINTER(OrTransfer), C.c_int, f64p, f64p]),
        "or_prolongate_add": (None, [C.POINTER(OrTransfer), C.c_int, f64p, f64p]),
        "or_hn_average": (None, [C.POINTER(OrHN), C.c_int, f64p]),
        "or_hn_distribute": (None, [C.POINTER(OrHN), C.c_int, f64p]),
        "or_hn_zero": (None, [C.POINTER(OrHN), C.c_int, f64p]),
        "or_dirichlet_zero": (None, [C.c_int, i32p, C.c_int, C.c_int, i32p, f64p]),
        "or_cginfo_init": (None, [C.POINTER(OrCGInfo), C.c_double, C.c_double, C.c_int]),
        "or_cginfo_reset": (None, [C.POINTER(OrCGInfo)]),
        "or_cginfo_check": (C.c_int, [C.POINTER(OrCGInfo), C.c_double, C.c_double]),
        "or_nlinfo_init": (None, [C.POINTER(OrNLInfo), C.c_double, C.c_double, C.c_int]),
        "or_nlinfo_check": (C.c_int, [C.POINTER(OrNLInfo), C.c_int, C.c_double, C.c_double]),
        "or_mg_create": (C.POINTER(OrMG), [C.c_int]),
        "or_mg_free": (None, [C.POINTER(OrMG)]),
        "or_mg_setup": (C.c_int, [C.POINTER(OrMG)]),
        "or_linear_mg": (None, [C.POINTER(OrMG), C.c_int, C.c_int, f64p, f64p, C.POINTER(OrCGInfo)]),
        "or_precondition": (None, [C.POINTER(OrMG), C.c_int, f64p, f64p]),
        "or_gmres_solve": (C.c_int, [C.POINTER(OrMG), C.c_int, C.c_int, f64p, f64p,
                                     C.POINTER(OrCGInfo)]),
        "or_newton_solve": (C.c_int, [C.POINTER(OrMG), C.c_int, C.c_int, f64p, f64p,
                                      C.POINTER(OrCGInfo), C.POINTER(OrNLInfo), C.c_int,
                                      C.POINTER(OrSolveReport)]),
        "or_num_threads": (C.c_int, []),
        "or_ilu_init": (C.c_int, [C.POINTER(OrIlu), C.POINTER(OrMatrix), i32p]),
        "or_ilu_solve": (None, [C.POINTER(OrIlu), f64p]),
        "or_ilu_free": (None, [C.POINTER(OrIlu)]),
        "or_set_parallel_vectors": (None, [C.c_int]),
        "or_dJdk_cells": (None, [C.c_int, f64p, C.c_int64, i32p, f64p, f64p, f64p]),
        "or_vanka_create": (C.c_void_p, [C.POINTER(OrMatrix), C.c_int, C.c_int, i32p]),
        "or_vanka_free": (None, [C.c_void_p]),
        "or_vanka_solve": (None, [C.c_void_p, f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def ptr(a, t=f64p):
    return a.ctypes.data_as(t)


class Matrix:
    """Oracle view of a block-CSR matrix (keeps the numpy arrays alive)."""

    def __init__(self, rowptr, col, val, ncomp=1):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rowptr) - 1
        self.ncomp = ncomp
        self.s = OrMatrix()
        rc = lib().or_matrix_init(C.byref(self.s), self.n, ncomp, ptr(self.rowptr, i32p),
                                  ptr(self.col, i32p), ptr(self.val))
        if rc:
            raise ValueError(f"or_matrix_init failed ({rc})")

    def __del__(self):
        try:
            lib().or_matrix_free(C.byref(self.s))
        except Exception:
            pass

    def vmult(self, y, x, s=1.0):
        lib().or_vmult(C.byref(self.s), ptr(y), ptr(x), s)

    def vmult_transpose(self, y, x, s=1.0):
        lib().or_vmult_transpose(C.byref(self.s), ptr(y), ptr(x), s)

    def residual(self, h, x, b):
        lib().or_residual(C.byref(self.s), ptr(h), ptr(x), ptr(b))

    def jacobi(self, h):
        lib().or_jacobi(C.byref(self.s), ptr(h))


class Ilu:
    """SparseBlockIlu in the node order perm (None: identity)"""

    def __init__(self, A, perm=None):
        self.A = A
        self.perm = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
        self.s = OrIlu()
        rc = lib().or_ilu_init(C.byref(self.s), C.byref(A.s),
                               ptr(self.perm, i32p) if self.perm is not None else None)
        if rc:
            raise ValueError(f"or_ilu_init failed ({rc})")
        n, B = A.n, A.ncomp
        nnz = int(A.rowptr[-1])
        self.rowptr = np.ctypeslib.as_array(self.s.rowptr, (n + 1,)).copy()
        self.col = np.ctypeslib.as_array(self.s.col, (max(nnz, 1),))[:nnz].copy()
        self.diag = np.ctypeslib.as_array(self.s.diag, (max(n, 1),))[:n].copy()
        self.val = np.ctypeslib.as_array(self.s.val, (max(nnz, 1) * B * B,))[:nnz * B * B].copy()

    def solve(self, x):
        lib().or_ilu_solve(C.byref(self.s), ptr(x))

    def __del__(self):
        try:
            lib().or_ilu_free(C.byref(self.s))
        except Exception:
            pass


class Transfer:
    def __init__(self, ncoarse, c2f, zweier, vierer, achter):
        self.keep = [np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
                     for a in (c2f, zweier, vierer, achter)]
        c2f, z, v, a = self.keep
        self.p = lib().or_transfer_create(ncoarse, ptr(c2f, i32p), len(z) // 3, ptr(z, i32p),
                                          len(v) // 5, ptr(v, i32p), len(a) // 9, ptr(a, i32p))
        self.ncoarse = ncoarse
        self.nfine = self.p.contents.nfine

    def __del__(self):
        try:
            lib().or_transfer_free(self.p)
        except Exception:
            pass

    def restrict_zero(self, uc, uf, ncomp=1):
        lib().or_restrict_zero(self.p, ncomp, ptr(uc), ptr(uf))

    def prolongate_add(self, uf, uc, ncomp=1):
        lib().or_prolongate_add(self.p, ncomp, ptr(uf), ptr(uc))

    def export(self):
        nnz = lib().or_transfer_nnz(self.p)
        crow = np.zeros(self.ncoarse + 1, np.int32)
        frow = np.zeros(self.nfine + 1, np.int32)
        ccol = np.zeros(nnz, np.int32)
        fcol = np.zeros(nnz, np.int32)
        cval = np.zeros(nnz)
        fval = np.zeros(nnz)
        lib().or_transfer_export(self.p, ptr(crow, i32p), ptr(ccol, i32p), ptr(cval),
                                 ptr(frow, i32p), ptr(fcol, i32p), ptr(fval))
        return crow, ccol, cval, frow, fcol, fval


class HN:
    def __init__(self, kind, edges, faces):
        self.edges = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 4)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 10)
        self.s = OrHN(kind, len(self.edges), len(self.faces), ptr(self.edges, i32p),
                      ptr(self.faces, i32p))

    def average(self, u, ncomp=1):
        lib().or_hn_average(C.byref(self.s), ncomp, ptr(u))

    def distribute(self, u, ncomp=1):
        lib().or_hn_distribute(C.byref(self.s), ncomp, ptr(u))

    def zero(self, u, ncomp=1):
        lib().or_hn_zero(C.byref(self.s), ncomp, ptr(u))


def cginfo(tol, globaltol, maxiter):
    info = OrCGInfo()
    lib().or_cginfo_init(C.byref(info), tol, globaltol, maxiter)
    return info


def nlinfo(tol, globaltol, maxiter):
    info = OrNLInfo()
    lib().or_nlinfo_init(C.byref(info), tol, globaltol, maxiter)
    return info


class MG:
    """Oracle multigrid hierarchy over a gascoignelib_amd.problem.Hierarchy-like
    object (solver numbering: level 0 = coarsest)."""

    def __init__(self, hier, *, mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8,
                 mgomega=1.0, smoother="jacobi", ndirect=1, patches=None, ilu_orders=None):
        nl = hier.nlevels
        self.hier = hier
        self.p = lib().or_mg_create(nl)
        M = self.p.contents
        M.mgtype = mgtype.encode()
        M.iterpre, M.iterpost, M.iterexact = iterpre, iterpost, iterexact
        M.omega, M.mgomega = omega, mgomega
        M.smoother = {"jacobi": 0, "richardson": 1, "none": 2, "vanka": 3, "ilu": 4}[smoother]
        self.vankas = []
        self.ilus = []
        self.keep = []
        self.mats, self.transfers, self.hns = [], [], []
        for l in range(nl):
            lev = hier.level(l)
            A = Matrix(lev.rowptr, lev.col, lev.val, lev.ncomp)
            self.mats.append(A)
            L = M.L[l]
            L.A = A.s
            if l > 0:
                T = Transfer(hier.level(l - 1).n, *lev.transfer)
                self.transfers.append(T)
                L.T = T.p
            hn = HN(lev.hn_kind, lev.hn_edges, lev.hn_faces)
            self.hns.append(hn)
            L.hn = hn.s
            d = np.ascontiguousarray(lev.dirichlet, dtype=np.int32)
            dc = np.ascontiguousarray(lev.dirichlet_comps, dtype=np.int32)
            self.keep += [d, dc]
            L.ndir, L.dir = len(d), ptr(d, i32p)
            L.ndircomp, L.dircomp = len(dc), ptr(dc, i32p)
            L.direct = 1 if lev.n < ndirect else 0
            if smoother == "vanka" and patches is not None and patches[l] is not None:
                V = Vanka(A, patches[l])
                self.vankas.append(V)
                L.vanka = V.p
            if smoother == "ilu":
                I = Ilu(A, None if ilu_orders is None else ilu_orders[l])
                self.ilus.append(I)
                L.ilu = C.cast(C.pointer(I.s), C.c_void_p)
        if lib().or_mg_setup(self.p):
            raise RuntimeError("or_mg_setup failed")

    def __del__(self):
        try:
            lib().or_mg_free(self.p)
        except Exception:
            pass

    @property
    def mgsteps(self):
        return self.p.contents.mgsteps

    def precondition(self, x, y, coarselevel=0):
        lib().or_precondition(self.p, coarselevel, ptr(x), ptr(y))

    def linear_mg(self, u, f, info, finelevel=None, coarselevel=0):
        fl = self.hier.nlevels - 1 if finelevel is None else finelevel
        lib().or_linear_mg(self.p, fl, coarselevel, ptr(u), ptr(f), C.byref(info))

    def gmres(self, x, b, info, vmax=10, coarselevel=0):
        return lib().or_gmres_solve(self.p, coarselevel, vmax, ptr(x), ptr(b), C.byref(info))

    def newton(self, u, f, linfo, nl, vmax=10, coarselevel=0, norm_kind=0):
        rep = OrSolveReport()
        rc = lib().or_newton_solve(self.p, coarselevel, vmax, ptr(u), ptr(f), C.byref(linfo),
                                   C.byref(nl), norm_kind, C.byref(rep))
        return rc, rep


def dJdk_cells(dim, coords, cells, z, u):
    """oracle per-cell coefficient gradient -z^T K^c u"""
    xyz = np.zeros((len(coords), 3))
    xyz[:, : np.asarray(coords).shape[1]] = coords
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    g = np.zeros(len(cells))
    lib().or_dJdk_cells(dim, ptr(np.ascontiguousarray(xyz)), len(cells), ptr(cells, i32p),
                        ptr(np.ascontiguousarray(z, dtype=np.float64)),
                        ptr(np.ascontiguousarray(u, dtype=np.float64)), ptr(g))
    return g


class Vanka:
    """oracle Vanka smoother over a Matrix and a (npatches, psize) node array"""

    def __init__(self, A, patches):
        self.A = A
        self.nodes = np.ascontiguousarray(patches, dtype=np.int32)
        self.p = lib().or_vanka_create(C.byref(A.s), self.nodes.shape[0], self.nodes.shape[1],
                                       ptr(self.nodes, i32p))

    def solve(self, x):
        lib().or_vanka_solve(self.p, ptr(x))

    def __del__(self):
        try:
            lib().or_vanka_free(self.p)
        except Exception:
            pass
