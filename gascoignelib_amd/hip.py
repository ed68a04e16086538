"""Python mirror of the reference GPU path's API surface on top of the C ABI.

Class and method names follow the reference's cuda/ backend so that code and
tests read like it:
  CudaVectorInterface      cuda/cudavectorinterface.h:26-77
  CudaCSRMatrixInterface   cuda/cudacsrmatrixinterface.h:40-159
  CudaMgInterpolatorNested cuda/cudamgintepolatornested.h:17-43
  CudaMultiLevelSolver     cuda/cudamultilevelsolver.h:15-45 (with the
                           StdMultiLevelSolver LinearMg / precondition / Gmres /
                           Solve entry points it inherits)
Every call goes to libgascoigne_hip.so; nothing here computes on the CPU.
"""
import ctypes as C

import numpy as np

from . import native as N
from .native import check, ptr


class Context:
    """Device, stream and reduction scratch (one per solver)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(N.hip().gh_ctx_create(device, C.byref(self.h)))

    def sync(self):
        check(N.hip().gh_ctx_sync(self.h))

    @property
    def stream(self):
        return N.hip().gh_ctx_stream(self.h)

    def close(self):
        if self.h:
            N.hip().gh_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CudaVectorInterface:
    def __init__(self, ctx, n, n_comp=1, prec="f64"):
        self.ctx = ctx
        self.n = int(n)
        self.n_comp = int(n_comp)
        self.prec = prec
        self.h = C.c_void_p()
        check(N.hip().gh_vec_create(ctx.h, self.n, self.n_comp, N.PREC[prec], C.byref(self.h)))

    @classmethod
    def from_global(cls, ctx, gv, n_comp=1, prec="f64"):
        gv = np.ascontiguousarray(gv, dtype=np.float64)
        v = cls(ctx, len(gv) // n_comp, n_comp, prec)
        v.assign(gv)
        return v

    def assign(self, gv):
        gv = np.ascontiguousarray(gv, dtype=np.float64)
        if gv.size != self.n * self.n_comp:
            raise ValueError("size mismatch")
        check(N.hip().gh_vec_upload(self.ctx.h, self.h, ptr(gv)))

    def copy_back(self):
        out = np.empty(self.n * self.n_comp)
        check(N.hip().gh_vec_download(self.ctx.h, self.h, ptr(out)))
        return out

    def zero(self):
        check(N.hip().gh_vec_zero(self.ctx.h, self.h))

    def copy_from(self, other):
        check(N.hip().gh_vec_copy(self.ctx.h, self.h, other.h))

    def equ(self, s, src):
        check(N.hip().gh_vec_equ(self.ctx.h, self.h, s, src.h))

    def add(self, d, y):
        check(N.hip().gh_vec_add(self.ctx.h, self.h, d, y.h))

    def sadd(self, a, d, y):
        check(N.hip().gh_vec_sadd(self.ctx.h, self.h, a, d, y.h))

    def scal(self, d):
        check(N.hip().gh_vec_scal(self.ctx.h, self.h, d))

    def norm(self):
        out = C.c_double()
        check(N.hip().gh_vec_norm(self.ctx.h, self.h, C.byref(out)))
        return out.value

    def norm_l8(self):
        out = C.c_double()
        check(N.hip().gh_vec_norm_l8(self.ctx.h, self.h, C.byref(out)))
        return out.value

    def dot(self, y):
        out = C.c_double()
        check(N.hip().gh_vec_dot(self.ctx.h, self.h, y.h, C.byref(out)))
        return out.value

    def free(self):
        if self.h:
            N.hip().gh_vec_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CudaCSRMatrixInterface:
    def __init__(self, ctx, rowptr, col, val, n_comp=1, prec="f64"):
        self.ctx = ctx
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        self._n = len(rowptr) - 1
        self.n_comp = n_comp
        self.prec = prec
        self.h = C.c_void_p()
        check(N.hip().gh_csr_create(ctx.h, self._n, n_comp, ptr(rowptr, N.i32p),
                                    ptr(col, N.i32p), ptr(val), N.PREC[prec], C.byref(self.h)))

    @classmethod
    def host_only(cls, ctx, rowptr, col, val, n_comp=1):
        """An FP64 matrix kept on the host only (gh_csr_create_host): a level
        of an ordered solver, which builds its own device matrices; device
        operations on it raise."""
        m = cls.__new__(cls)
        m.ctx = ctx
        m._keep = (np.ascontiguousarray(rowptr, dtype=np.int32),
                   np.ascontiguousarray(col, dtype=np.int32),
                   np.ascontiguousarray(val, dtype=np.float64))
        m._n = len(m._keep[0]) - 1
        m.n_comp = n_comp
        m.prec = "f64"
        m.h = C.c_void_p()
        check(N.hip().gh_csr_create_host(ctx.h, m._n, n_comp, ptr(m._keep[0], N.i32p),
                                         ptr(m._keep[1], N.i32p), ptr(m._keep[2]),
                                         C.byref(m.h)))
        m._keep = None
        return m

    @classmethod
    def view(cls, ctx, handle, n_comp, owner):
        """A matrix owned by another object (`owner` is kept alive; the handle
        is never destroyed through the view)."""
        v = cls.__new__(cls)
        v.ctx = ctx
        v.h = C.c_void_p(handle)
        v._n = int(N.hip().gh_csr_n(v.h))
        v.n_comp = n_comp
        v.prec = {0: "f64", 1: "f32", 2: "f16"}.get(int(N.hip().gh_csr_prec(v.h)), "f64")
        v._owner = owner
        return v

    def n(self):
        return self._n

    def n_entries(self):
        return N.hip().gh_csr_nnz(self.h)

    def layout(self):
        """Device layout: entries stored, padded slots, slices, slices with
        16-bit column offsets, slices in pattern mode, pattern-table entries,
        column bytes a sweep reads, whether rows run in a permuted order."""
        out = np.zeros(8, np.int64)
        N.check(N.hip().gh_csr_layout(self.h, N.ptr(out, N.i64p)))
        return dict(stored=int(out[0]), slots=int(out[1]), slices=int(out[2]),
                    narrow=int(out[3]), patterned=int(out[4]), table=int(out[5]),
                    column_bytes=int(out[6]), permuted=bool(out[7]))

    def vmult(self, y, x, d=1.0, b=0.0):
        """y = d*A*x + b*y"""
        check(N.hip().gh_spmv(self.ctx.h, self.h, y.h, x.h, d, b))

    def vmult_transpose(self, y, x, d=1.0, b=0.0):
        check(N.hip().gh_spmv_t(self.ctx.h, self.h, y.h, x.h, d, b))

    def residual(self, r, x, b):
        check(N.hip().gh_residual(self.ctx.h, self.h, r.h, x.h, b.h))

    def smooth(self, x, b, h, niter, omega, smoother="jacobi"):
        check(N.hip().gh_smooth(self.ctx.h, self.h, x.h, b.h, h.h, niter, omega,
                                {"jacobi": 0, "richardson": 1}[smoother]))

    def jacobi(self, h):
        check(N.hip().gh_jacobi(self.ctx.h, self.h, h.h))

    def values(self):
        """the values on the device, in CSR order, as FP64"""
        nnz = N.hip().gh_csr_nnz(self.h)
        out = np.empty(nnz * self.n_comp * self.n_comp)
        check(N.hip().gh_csr_download_values(self.ctx.h, self.h, ptr(out)))
        return out

    def jacobi_noalloc(self, h):
        """h = D^-1 h through the fixed per-GPU scratch (no device allocation)"""
        check(N.hip().gh_jacobi_noalloc(self.ctx.h, self.h, h.h))

    def free(self):
        if getattr(self, "_owner", None) is not None:  # a view: the owner frees
            self.h = C.c_void_p()
            self._owner = None
            return
        if self.h:
            N.hip().gh_csr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CudaMgInterpolatorNested:
    def __init__(self, ctx, ncoarse, c2f, zweier, vierer, achter):
        self.ctx = ctx
        arrs = [np.ascontiguousarray(a, dtype=np.int32).reshape(-1) for a in
                (c2f, zweier, vierer, achter)]
        c2f, z, v, a = arrs
        self.h = C.c_void_p()
        check(N.hip().gh_transfer_create(ctx.h, int(ncoarse), ptr(c2f, N.i32p), len(z) // 3,
                                         ptr(z, N.i32p), len(v) // 5, ptr(v, N.i32p),
                                         len(a) // 9, ptr(a, N.i32p), C.byref(self.h)))

    def restrict_zero(self, uc, uf):
        check(N.hip().gh_restrict_zero(self.ctx.h, self.h, uc.h, uf.h))

    def prolongate_add(self, uf, uc):
        check(N.hip().gh_prolongate_add(self.ctx.h, self.h, uf.h, uc.h))

    def free(self):
        if self.h:
            N.hip().gh_transfer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HNStructure:
    """Hanging-node constraints (HNStructureQ12d/Q13d/Q22d/Q23d semantics)."""

    def __init__(self, ctx, kind, n, edges, faces):
        self.ctx = ctx
        e = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 4)
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 10)
        self.h = C.c_void_p()
        check(N.hip().gh_hn_create(ctx.h, int(kind), int(n), len(e), ptr(e, N.i32p), len(f),
                                   ptr(f, N.i32p), C.byref(self.h)))

    def average(self, x):
        check(N.hip().gh_hn_average(self.ctx.h, self.h, x.h))

    def distribute(self, x):
        check(N.hip().gh_hn_distribute(self.ctx.h, self.h, x.h))

    def zero(self, x):
        check(N.hip().gh_hn_zero(self.ctx.h, self.h, x.h))

    def free(self):
        if self.h:
            N.hip().gh_hn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DirichletMask:
    def __init__(self, ctx, n, ncomp, nodes, comps=(0,)):
        self.ctx = ctx
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        comps = np.ascontiguousarray(comps, dtype=np.int32)
        self.h = C.c_void_p()
        check(N.hip().gh_mask_create(ctx.h, int(n), int(ncomp), len(nodes), ptr(nodes, N.i32p),
                                     len(comps), ptr(comps, N.i32p), C.byref(self.h)))

    def zero(self, x):
        check(N.hip().gh_mask_zero(self.ctx.h, self.h, x.h))

    def free(self):
        if self.h:
            N.hip().gh_mask_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CGInfo:
    def __init__(self, tol=1e-6, globaltol=1e-14, maxiter=100):
        self.s = N.GhCGInfo()
        N.hip().gh_cginfo_init(C.byref(self.s), tol, globaltol, maxiter)

    def reset(self):
        N.hip().gh_cginfo_reset(C.byref(self.s))

    def check(self, res, cor=0.0):
        return N.hip().gh_cginfo_check(C.byref(self.s), res, cor)

    @property
    def status(self):
        return N.CG_STATUS[self.s.status]

    def __getattr__(self, k):
        return getattr(self.__dict__["s"], k)


class NLInfo:
    def __init__(self, tol=1e-4, globaltol=1e-12, maxiter=10):
        self.s = N.GhNLInfo()
        N.hip().gh_nlinfo_init(C.byref(self.s), tol, globaltol, maxiter)

    @property
    def status(self):
        return N.NL_STATUS[self.s.status]

    def __getattr__(self, k):
        return getattr(self.__dict__["s"], k)


class CudaMultiLevelSolver:
    """Device multigrid hierarchy built from a gascoignelib_amd.problem.Hierarchy.

    Parameters follow the reference's MultiLevelSolver / Solver blocks:
    mgtype, iterpre, iterpost, iterexact, omega, mgomega, linear_smooth
    ("jacobi", "richardson", "vanka" with `patches`: per level an
    (npatches, psize) node array or None, or "ilu": ILU(0) factors on every
    level, ILU row i = node ilu_orders[l][i], the identity without
    `ilu_orders`),
    ndirect, prec_maxiter; `prec` selects the V-cycle storage ("f64", "f32"
    or "f16" matrix values with FP32 arithmetic); `krylov` the outer solver of
    Gmres/Solve: "gmres" (left-preconditioned, as GMRES::solve) or "fgmres"
    (flexible, right-preconditioned); `adjoint` makes Gmres/Solve/LinearMg
    solve with the transposed matrices (A^T z = g)."""

    def __init__(self, ctx, hier, mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8,
                 mgomega=1.0, linear_smooth="jacobi", ndirect=1, prec_maxiter=1, prec="f64",
                 use_graph=False, order=None, krylov="gmres", adjoint=False, patches=None,
                 ilu_orders=None):
        """order: None (keep the hierarchy's numbering), "lexicographic" (sort
        each level's nodes by their coordinates) or a list of per-level
        permutations (internal position -> node)."""
        self.ctx = ctx
        self.hier = hier
        nl = hier.nlevels
        self.A, self.T, self.H, self.M = [], [], [], []
        self._fine_device = None
        self._ordered = order is not None
        for l in range(nl):
            lev = hier.level(l)
            if order is None:
                self.A.append(CudaCSRMatrixInterface(ctx, lev.rowptr, lev.col, lev.val, lev.ncomp,
                                                     "f64"))
            else:
                # the ordered solver builds its own device matrices from these
                self.A.append(CudaCSRMatrixInterface.host_only(ctx, lev.rowptr, lev.col, lev.val,
                                                               lev.ncomp))
            if l > 0:
                self.T.append(CudaMgInterpolatorNested(ctx, hier.level(l - 1).n, *lev.transfer))
            else:
                self.T.append(None)
            if len(lev.hn_edges) or len(lev.hn_faces):
                self.H.append(HNStructure(ctx, lev.hn_kind, lev.n, lev.hn_edges, lev.hn_faces))
            else:
                self.H.append(None)
            if len(lev.dirichlet):
                self.M.append(DirichletMask(ctx, lev.n, lev.ncomp, lev.dirichlet, lev.dirichlet_comps))
            else:
                self.M.append(None)
        p = N.GhMgParams()
        N.hip().gh_mg_params_default(C.byref(p))
        p.mgtype = mgtype.encode()
        p.iterpre, p.iterpost, p.iterexact = iterpre, iterpost, iterexact
        p.omega, p.mgomega = omega, mgomega
        p.smoother = {"jacobi": 0, "richardson": 1, "vanka": 2, "ilu": 3}[linear_smooth]
        p.ndirect = ndirect
        p.prec_maxiter = prec_maxiter
        p.prec = N.PREC[prec]
        p.use_graph = int(use_graph)
        p.krylov = {"gmres": 0, "fgmres": 1}[krylov]
        p.adjoint = int(adjoint)
        self.params = p

        def arr(objs):
            a = (C.c_void_p * nl)()
            for i, o in enumerate(objs):
                a[i] = o.h if o is not None else None
            return a

        self._arrays = [arr(self.A), arr(self.T), arr(self.H), arr(self.M)]
        self.h = C.c_void_p()
        if order is None:
            check(N.hip().gh_mg_create(ctx.h, nl, self._arrays[0], self._arrays[1],
                                       self._arrays[2], self._arrays[3], C.byref(p),
                                       C.byref(self.h)))
        else:
            if isinstance(order, str):
                from .problem import lexicographic_order
                if order != "lexicographic":
                    raise ValueError(f"unknown order {order!r}")
                order = [lexicographic_order(hier.level(l).coords) for l in range(nl)]
            self._order = [np.ascontiguousarray(o, dtype=np.int32) for o in order]
            oarr = (N.i32p * nl)(*[ptr(o, N.i32p) for o in self._order])
            check(N.hip().gh_mg_create_ordered(ctx.h, nl, self._arrays[0], self._arrays[1],
                                               self._arrays[2], self._arrays[3], oarr,
                                               C.byref(p), C.byref(self.h)))
        if patches is not None:
            for l, pt in enumerate(patches):
                if pt is None:
                    continue
                pt = np.ascontiguousarray(pt, dtype=np.int32)
                check(N.hip().gh_mg_set_vanka(self.h, l, pt.shape[0], pt.shape[1],
                                              ptr(pt, N.i32p)))
        if linear_smooth == "ilu":
            for l in range(nl):
                po = None
                if ilu_orders is not None and ilu_orders[l] is not None:
                    po = np.ascontiguousarray(ilu_orders[l], dtype=np.int32)
                check(N.hip().gh_mg_set_ilu(self.h, l, None if po is None else ptr(po, N.i32p)))

    @property
    def fine_matrix(self):
        """The finest-level matrix in the caller's numbering, on the device
        (for an ordered solver made on first use: its levels are host-only)."""
        if self.A[-1].h and not self._ordered:
            return self.A[-1]
        if self._fine_device is None:
            lev = self.hier.level(self.hier.nlevels - 1)
            self._fine_device = CudaCSRMatrixInterface(self.ctx, lev.rowptr, lev.col, lev.val,
                                                       lev.ncomp, "f64")
        return self._fine_device

    def level_matrix(self, level=None, fp64=True):
        """The matrix the solver runs on `level` (finest by default), in its
        internal numbering: a view owned by the solver."""
        l = self.hier.nlevels - 1 if level is None else level
        h = N.hip().gh_mg_level_matrix(self.h, l, int(bool(fp64)))
        if not h:
            raise ValueError(f"no matrix on level {l}")
        return CudaCSRMatrixInterface.view(self.ctx, h, self.hier.level(l).ncomp, self)

    @property
    def fine_mask(self):
        return self.M[-1]

    @property
    def fine_hn(self):
        return self.H[-1]

    @property
    def mgsteps(self):
        return N.hip().gh_mg_mgsteps(self.h)

    def new_vector(self, level=None, prec="f64"):
        lev = self.hier.level(self.hier.nlevels - 1 if level is None else level)
        return CudaVectorInterface(self.ctx, lev.n, lev.ncomp, prec)

    def LinearMg(self, u, f, info, finelevel=None, coarselevel=0):
        fl = self.hier.nlevels - 1 if finelevel is None else finelevel
        check(N.hip().gh_mg_linear(self.ctx.h, self.h, fl, coarselevel, u.h, f.h, C.byref(info.s)))

    def precondition(self, x, y, x_is_zero=False, coarselevel=0):
        check(N.hip().gh_mg_precondition(self.ctx.h, self.h, coarselevel, x.h, y.h, int(x_is_zero)))

    def Gmres(self, x, b, info, vmax=10, coarselevel=0):
        reached = C.c_int()
        m = self.fine_mask
        check(N.hip().gh_gmres_solve(self.ctx.h, self.h, self.A[-1].h, m.h if m else None,
                                     coarselevel, vmax, x.h, b.h, C.byref(info.s), C.byref(reached)))
        return reached.value

    def Solve(self, u, f, linfo, nlinfo, vmax=10, coarselevel=0, norm_kind=0):
        rep = N.GhSolveReport()
        m, hn = self.fine_mask, self.fine_hn
        check(N.hip().gh_newton_solve(self.ctx.h, self.h, self.A[-1].h, m.h if m else None,
                                      hn.h if hn else None, coarselevel, vmax, u.h, f.h,
                                      C.byref(linfo.s), C.byref(nlinfo.s), norm_kind,
                                      C.byref(rep)))
        return rep

    def assemble_q1(self, cells, k=None):
        """AssembleMatrix + ComputeIlu on the device: every level's operator
        re-formed from its cells (CoefficientCells per level, caller numbering)
        and per-cell coefficients (CudaVectorInterface per level, or None)"""
        nl = self.hier.nlevels
        carr = (C.c_void_p * nl)(*[c.h for c in cells])
        karr = None
        if k is not None:
            karr = (C.c_void_p * nl)(*[(v.h if v is not None else None) for v in k])
        check(N.hip().gh_mg_assemble_q1(self.ctx.h, self.h, carr, karr))

    def assemble_q2(self, patches, k=None):
        """The same for a hierarchy of triquadratic patch elements without
        hanging nodes: every level's operator re-formed from its patches
        (DwrPatches per level, caller numbering) and per-patch coefficients
        (CudaVectorInterface per level, or None)"""
        nl = self.hier.nlevels
        parr = (C.c_void_p * nl)(*[p.h for p in patches])
        karr = None
        if k is not None:
            karr = (C.c_void_p * nl)(*[(v.h if v is not None else None) for v in k])
        check(N.hip().gh_mg_assemble_q2(self.ctx.h, self.h, parr, karr))

    def free(self):
        if self.h:
            N.hip().gh_mg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CoefficientCells:
    """Q1 box cells of a mesh level for per-cell coefficient gradients
    (gh_cells_create / gh_dJdk_cells): dJdk(z, u, g) writes
    g_c = -z^T (dA/dk_c) u for the state u and the adjoint z."""

    def __init__(self, ctx, dim, coords, cells):
        self.ctx = ctx
        xyz = np.zeros((len(coords), 3))
        xyz[:, : np.asarray(coords).shape[1]] = coords
        self._xyz = np.ascontiguousarray(xyz)
        self._cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.ncells = len(self._cells)
        self.n = len(xyz)
        self.h = C.c_void_p()
        check(N.hip().gh_cells_create(ctx.h, dim, self.n, ptr(self._xyz), self.ncells,
                                      ptr(self._cells, N.i32p), C.byref(self.h)))

    def dJdk(self, z, u, g):
        check(N.hip().gh_dJdk_cells(self.ctx.h, self.h, z.h, u.h, g.h))

    def assemble(self, A, k=None, mask=None):
        """A <- Q1 operator of -div(k grad u) on these cells, formed on the
        device (k: CudaVectorInterface with one entry per cell, or None)"""
        check(N.hip().gh_assemble_q1(self.ctx.h, self.h, k.h if k is not None else None,
                                     mask.h if mask is not None else None, A.h))

    def form(self, r, u, f, k=None, mask=None):
        """r = f - A(u), cell by cell (CGDisc::Form), mask rows zero"""
        check(N.hip().gh_form_q1(self.ctx.h, self.h, k.h if k is not None else None,
                                 mask.h if mask is not None else None, u.h, f.h, r.h))

    def __del__(self):
        try:
            if self.h:
                N.hip().gh_cells_destroy(self.h)
        except Exception:
            pass


class DwrPatches:
    """Patches of a refined mesh level (gh_patches_create), for dual-weighted
    residual indicators (gh_dwr_estimate) and for the device assembly of
    triquadratic patch elements (gh_assemble_q2, gh_form_q2).
    estimate(u, z, f, k) returns the
    per-cell eta_K = int_K f w - k_K int_K grad u_h . grad w, w = I2 z_h - z_h,
    as a device vector of npatches * 2^dim values (each patch's cells in
    lexicographic order); their sum estimates J(u) - J(u_h)."""

    def __init__(self, ctx, dim, coords, patches):
        self.ctx = ctx
        xyz = np.zeros((len(coords), 3))
        xyz[:, : np.asarray(coords).shape[1]] = coords
        self._xyz = np.ascontiguousarray(xyz)
        self._nodes = np.ascontiguousarray(patches, dtype=np.int32)
        self.dim = dim
        self.npatches = len(self._nodes)
        self.ncells = self.npatches * 2 ** dim
        self.h = C.c_void_p()
        check(N.hip().gh_patches_create(ctx.h, dim, len(xyz), ptr(self._xyz), self.npatches,
                                        ptr(self._nodes, N.i32p), C.byref(self.h)))

    def assemble_q2(self, A, k=None, mask=None):
        """A <- the operator of -div(k grad u) on the triquadratic patch
        elements of these patches (CGDiscQ23d), formed on the device (k:
        CudaVectorInterface with one entry per patch, or None)"""
        check(N.hip().gh_assemble_q2(self.ctx.h, self.h, k.h if k is not None else None,
                                     mask.h if mask is not None else None, A.h))

    def form_q2(self, r, u, f, k=None, mask=None):
        """r = f - A(u), patch element by patch element (CGDisc::Form), mask
        rows zero"""
        check(N.hip().gh_form_q2(self.ctx.h, self.h, k.h if k is not None else None,
                                 mask.h if mask is not None else None, u.h, f.h, r.h))

    def estimate(self, u, z, f, k=None, eta=None):
        if eta is None:
            eta = CudaVectorInterface(self.ctx, self.ncells)
        check(N.hip().gh_dwr_estimate(self.ctx.h, self.h, u.h, z.h, f.h,
                                      k.h if k is not None else None, eta.h))
        return eta

    def __del__(self):
        try:
            if self.h:
                N.hip().gh_patches_destroy(self.h)
        except Exception:
            pass


def condense_q2(ctx, edges, faces, A, Ac, mask=None):
    """Ac <- T^T A T: the unconstrained FP64 Q2 operator A (DwrPatches.assemble_q2)
    condensed over the Q2 hanging nodes (edges (ne, 4) [h, near, middle, far],
    faces (nf, 10) [h, 9 parents]) on the device (gh_condense_q2); hanging rows
    identity rows, then the mask's rows and columns"""
    e = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 4)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 10)
    check(N.hip().gh_condense_q2(ctx.h, len(e), ptr(e, N.i32p), len(f), ptr(f, N.i32p), A.h,
                                 mask.h if mask is not None else None, Ac.h))


class SparseBlockIlu:
    """SparseBlockIlu: ILU(0) factors of an FP64 matrix, ILU row i = node
    perm[i]; solve(x) in place (FP64 or FP32 vector)."""

    def __init__(self, ctx, A, perm=None):
        self.ctx = ctx
        self._perm = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
        self.h = C.c_void_p()
        self.n, self.ncomp = A._n, A.n_comp
        check(N.hip().gh_ilu_create(ctx.h, A.h, None if self._perm is None else
                                    ptr(self._perm, N.i32p), C.byref(self.h)))

    def solve(self, x):
        check(N.hip().gh_ilu_solve(self.ctx.h, self.h, x.h))

    def levels(self):
        f, b = C.c_int32(), C.c_int32()
        check(N.hip().gh_ilu_levels(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def factors(self, nnz):
        """(rowptr, col, diag, val) of the factors in ILU numbering."""
        B = self.ncomp
        rp = np.empty(self.n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        dg = np.empty(self.n, np.int32)
        val = np.empty(nnz * B * B)
        check(N.hip().gh_ilu_download(self.ctx.h, self.h, ptr(rp, N.i32p), ptr(col, N.i32p),
                                      ptr(dg, N.i32p), ptr(val)))
        return rp, col, dg, val

    def __del__(self):
        try:
            if self.h:
                N.hip().gh_ilu_destroy(self.h)
        except Exception:
            pass


class VankaSmoother:
    """VankaSmoother: patch LU factors of an FP64 matrix; solve(x) in place."""

    def __init__(self, ctx, A, patches, prec="f64"):
        self.ctx = ctx
        self._nodes = np.ascontiguousarray(patches, dtype=np.int32)
        self.h = C.c_void_p()
        check(N.hip().gh_vanka_create(ctx.h, A.h, self._nodes.shape[0], self._nodes.shape[1],
                                      ptr(self._nodes, N.i32p), N.PREC[prec], C.byref(self.h)))

    def solve(self, x):
        check(N.hip().gh_vanka_solve(self.ctx.h, self.h, x.h))

    def __del__(self):
        try:
            if self.h:
                N.hip().gh_vanka_destroy(self.h)
        except Exception:
            pass
