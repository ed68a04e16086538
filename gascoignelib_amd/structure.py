"""Multigrid problems from the native structure producer (libgh_structure.so).

`uniform_hierarchy` refines a coarse .inp mesh uniformly and returns a
`problem.Hierarchy` with the Q1 matrices of -div(k grad u) (k = 1 unless a
per-cell coefficient is given for the finest level), the mesh-transfer maps
and the strong Dirichlet node lists of the given colours, in the form
Gascoigne's CPU side hands them to the GPU path.
"""
import ctypes as C
import os

import numpy as np

from . import native as N
from .native import check_s, ptr
from .problem import Hierarchy, Level

CONFIGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")


class Mesh:
    """Handle on a refined mesh hierarchy inside libgh_structure.so."""

    def __init__(self, inp, dim, prerefine, drop_coarse=0):
        path = inp if os.path.isabs(inp) else os.path.join(CONFIGS, inp)
        self.dim = dim
        self.h = C.c_void_p()
        check_s(N.structure().ghs_build(path.encode(), dim, prerefine, drop_coarse, C.byref(self.h)))
        self.nlevels = N.structure().ghs_nlevels(self.h)

    @classmethod
    def adaptive(cls, inp, prerefine, drop_coarse, nadapt, radius, centre):
        """3D hierarchy of `prerefine` uniform refinements and `nadapt`
        adaptive patch refinements inside a ball (ghs_build_adaptive)"""
        m = cls.__new__(cls)
        path = inp if os.path.isabs(inp) else os.path.join(CONFIGS, inp)
        m.dim = 3
        m.h = C.c_void_p()
        ctr = np.ascontiguousarray(centre, dtype=np.float64)
        check_s(N.structure().ghs_build_adaptive(path.encode(), prerefine, drop_coarse, nadapt,
                                                 float(radius), ptr(ctr), C.byref(m.h)))
        m.nlevels = N.structure().ghs_nlevels(m.h)
        return m

    @classmethod
    def from_gup(cls, path, drop_coarse=1):
        """3D hierarchy of a Gascoigne hierarchical mesh file (ghs_build_gup):
        the macro mesh and finest cells it holds, with the multigrid levels
        Gascoigne derives from them"""
        m = cls.__new__(cls)
        m.dim = 3
        m.h = C.c_void_p()
        check_s(N.structure().ghs_build_gup(os.fspath(path).encode(), drop_coarse, C.byref(m.h)))
        m.nlevels = N.structure().ghs_nlevels(m.h)
        return m

    def sizes(self, l, with_nnz=False):
        n, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        check_s(N.structure().ghs_level_sizes(self.h, l, C.byref(n), C.byref(c),
                                              C.byref(z) if with_nnz else None))
        return n.value, c.value, (z.value if with_nnz else None)

    def coords(self, l):
        n, _, _ = self.sizes(l)
        x = np.empty((n, 3))
        check_s(N.structure().ghs_level_coords(self.h, l, ptr(x)))
        return x[:, : self.dim].copy()

    def cells(self, l):
        _, nc, _ = self.sizes(l)
        npc = 8 if self.dim == 3 else 4
        c = np.empty((nc, npc), np.int32)
        check_s(N.structure().ghs_level_cells(self.h, l, ptr(c, N.i32p)))
        return c

    def structure(self, l):
        n, _, nnz = self.sizes(l, with_nnz=True)
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        check_s(N.structure().ghs_level_structure(self.h, l, ptr(rp, N.i32p), ptr(col, N.i32p)))
        return rp, col

    def dirichlet(self, l, colours=(1,)):
        cols = np.ascontiguousarray(colours, dtype=np.int32)
        cnt = C.c_int64()
        check_s(N.structure().ghs_level_dirichlet_count(self.h, l, len(cols), ptr(cols, N.i32p),
                                                        C.byref(cnt)))
        out = np.empty(cnt.value, np.int32)
        check_s(N.structure().ghs_level_dirichlet(self.h, l, len(cols), ptr(cols, N.i32p),
                                                  ptr(out, N.i32p)))
        return out

    def assemble(self, l, nnz, kcell=None, colours=(1,)):
        cols = np.ascontiguousarray(colours, dtype=np.int32)
        val = np.empty(nnz)
        kp = None
        if kcell is not None:
            kcell = np.ascontiguousarray(kcell, dtype=np.float64)
            kp = ptr(kcell)
        check_s(N.structure().ghs_level_assemble(self.h, l, kp, len(cols), ptr(cols, N.i32p),
                                                 ptr(val)))
        return val

    def transfer(self, l):
        nc, n2, n4, n8 = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check_s(N.structure().ghs_level_transfer_sizes(self.h, l, C.byref(nc), C.byref(n2),
                                                       C.byref(n4), C.byref(n8)))
        c2f = np.empty(nc.value, np.int32)
        z = np.empty((n2.value, 3), np.int32)
        v = np.empty((n4.value, 5), np.int32)
        a = np.empty((n8.value, 9), np.int32)
        check_s(N.structure().ghs_level_transfer(self.h, l, ptr(c2f, N.i32p), ptr(z, N.i32p),
                                                 ptr(v, N.i32p), ptr(a, N.i32p)))
        return c2f, z, v, a

    def patch_sizes(self, l):
        npat, nnz = C.c_int64(), C.c_int64()
        check_s(N.structure().ghs_level_patch_sizes(self.h, l, C.byref(npat), C.byref(nnz)))
        return npat.value, nnz.value

    def patches(self, l):
        """(npatches, 27) patch nodes, patch-local lexicographic order"""
        npat, _ = self.patch_sizes(l)
        out = np.empty((npat, 27 if self.dim == 3 else 9), np.int32)
        check_s(N.structure().ghs_level_patches(self.h, l, ptr(out, N.i32p)))
        return out

    def patch_structure(self, l):
        n, _, _ = self.sizes(l)
        _, nnz = self.patch_sizes(l)
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        check_s(N.structure().ghs_level_patch_structure(self.h, l, ptr(rp, N.i32p),
                                                        ptr(col, N.i32p)))
        return rp, col

    def assemble_stokes(self, l, nnz, visc, alpha, colours, dircomps, u0=None):
        """4x4 blocks of the Stokes-LPS operator on the patch structure; with
        u0 (n, 4) also the lifted residual -A u0 (zero in constrained entries)."""
        cols = np.ascontiguousarray(colours, dtype=np.int32)
        mask = 0
        for c in dircomps:
            mask |= 1 << int(c)
        val = np.empty(nnz * 16)
        rhs = None
        if u0 is not None:
            u0 = np.ascontiguousarray(u0, dtype=np.float64)
            rhs = np.empty(u0.size)
        check_s(N.structure().ghs_level_assemble_stokes(
            self.h, l, float(visc), float(alpha), len(cols), ptr(cols, N.i32p), mask,
            ptr(u0) if u0 is not None else None, ptr(rhs) if rhs is not None else None, ptr(val)))
        return val, rhs

    def assemble_q2(self, l, nnz, colours=(1,)):
        """Laplace on the triquadratic patch element, values on the patch structure"""
        cols = np.ascontiguousarray(colours, dtype=np.int32)
        val = np.empty(nnz)
        check_s(N.structure().ghs_level_assemble_q2(self.h, l, len(cols), ptr(cols, N.i32p),
                                                    ptr(val)))
        return val

    def hn(self, l):
        """Q2 hanging-node constraints: (edges (ne, 4), faces (nf, 10))"""
        ne, nf = C.c_int64(), C.c_int64()
        check_s(N.structure().ghs_level_hn_sizes(self.h, l, C.byref(ne), C.byref(nf)))
        e = np.empty((ne.value, 4), np.int32)
        f = np.empty((nf.value, 10), np.int32)
        check_s(N.structure().ghs_level_hn(self.h, l, ptr(e, N.i32p), ptr(f, N.i32p)))
        return e, f

    def q2_condensed(self, l, colours=(1,)):
        """Q2 Laplace condensed over the hanging nodes, Dirichlet rows of the
        given colours: (rowptr, col, val)"""
        cols = np.ascontiguousarray(colours, dtype=np.int32)
        nnz = C.c_int64()
        check_s(N.structure().ghs_level_q2_condensed_nnz(self.h, l, len(cols), ptr(cols, N.i32p),
                                                         C.byref(nnz)))
        n, _, _ = self.sizes(l)
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value)
        check_s(N.structure().ghs_level_q2_condensed(self.h, l, ptr(rp, N.i32p), ptr(col, N.i32p),
                                                     ptr(val)))
        return rp, col, val

    def face_mass(self, l, colour):
        n, _, _ = self.sizes(l)
        m = np.empty(n)
        check_s(N.structure().ghs_level_face_mass(self.h, l, int(colour), ptr(m)))
        return m

    def __del__(self):
        try:
            if self.h:
                N.structure().ghs_destroy(self.h)
        except Exception:
            pass


def colour_order(rowptr, col):
    """ILU node order with few triangular-solve levels (ghs_colour_order):
    the nodes grouped by a greedy colouring of the matrix graph.  Returns
    (perm, ncolours); ILU row i is node perm[i]."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    n = len(rowptr) - 1
    perm = np.empty(n, np.int32)
    nc = C.c_int32()
    check_s(N.structure().ghs_colour_order(n, ptr(rowptr, N.i32p), ptr(col, N.i32p),
                                           ptr(perm, N.i32p), C.byref(nc)))
    return perm, nc.value


def coarse_coefficients(k_fine, cells_per_coarse=8):
    """Per-cell coefficients of the next coarser level: the mean over the
    children (children of coarse cell c are fine cells 8c..8c+7 in 3D)."""
    return np.asarray(k_fine).reshape(-1, cells_per_coarse).mean(axis=1)


def uniform_hierarchy(inp="cube.inp", dim=3, prerefine=7, drop_coarse=1, colours=(1,),
                      kcell=None, name=None, keep_coords=True):
    """Q1 hierarchy of a uniformly refined mesh.  drop_coarse removes the
    coarsest meshes from the hierarchy (Gascoigne keeps nlevels - patchdepth
    levels: the 1-cell mesh of a one-hex .inp is not a multigrid level)."""
    M = Mesh(inp, dim, prerefine, drop_coarse)
    return assemble_hierarchy(M, kcell, colours, name or f"{inp}:{prerefine}", keep_coords)


def assemble_hierarchy(M, kcell=None, colours=(1,), name="", keep_coords=True):
    """Matrices, transfers and Dirichlet lists of every level of the mesh M;
    kcell (one coefficient per finest cell) is averaged onto coarser cells."""
    nl = M.nlevels
    dim = M.dim
    ks = [None] * nl
    if kcell is not None:
        ks[-1] = np.asarray(kcell, dtype=np.float64)
        for l in range(nl - 1, 0, -1):
            ks[l - 1] = coarse_coefficients(ks[l], 8 if dim == 3 else 4)
    levels = []
    for l in range(nl):
        rp, col = M.structure(l)
        val = M.assemble(l, len(col), ks[l], colours)
        lev = Level(n=len(rp) - 1, ncomp=1, rowptr=rp, col=col, val=val,
                    transfer=M.transfer(l) if l > 0 else None, hn_kind=13 if dim == 3 else 12,
                    dirichlet=M.dirichlet(l, colours), dirichlet_comps=np.zeros(1, np.int32),
                    coords=M.coords(l) if keep_coords else None)
        levels.append(lev)
    h = Hierarchy(levels, dim, name)
    h.mesh = M
    h.kcell = ks
    return h


def inflow_profile(coords, amplitude):
    """Parabolic inflow u = amplitude * 16 y(1-y) z(1-z) on the unit cube's
    x = 0 face (zero on its edges)."""
    y, z = coords[:, 1], coords[:, 2]
    return amplitude * 16.0 * y * (1.0 - y) * z * (1.0 - z)


def inflow_amplitude(seed):
    """Seeded inflow amplitude, uniform in [0.5, 1.5) (SplitMix64 position 0)."""
    from .problem import splitmix64_uniform
    return 1.0 + 0.5 * float(splitmix64_uniform(seed, 1)[0])


def stokes_hierarchy(prerefine, seed=0, visc=1.0, alpha=0.1, inp="channel.inp", drop_coarse=1,
                     name=None, keep_coords=True, with_full=False):
    """Stokes-LPS channel on the unit cube (NavierStokesLps<3> at U = 0,
    CGDiscQ13dLps): components (p, u, v, w), 4x4 blocks on the patch
    structure, velocity Dirichlet on colours 1 (inflow, x = 0, parabolic
    profile of seeded amplitude) and 2 (walls, zero), do-nothing outflow on
    colour 3 (x = 1).  The finest level carries the Newton residual of the
    lifted inflow data: h.rhs (n*4), h.u0 (n, 4); per level h.patches
    (npatches, 27) for the Vanka smoother, h.mass_in / h.mass_out the
    boundary masses of the inflow and outflow faces on the finest level;
    with_full: h.full_val, the finest operator without Dirichlet conditions
    (the lifting and the adjoint gradient with respect to boundary data)."""
    M = Mesh(inp, 3, prerefine, drop_coarse)
    nl = M.nlevels
    colours, comps = (1, 2), (1, 2, 3)
    amp = inflow_amplitude(seed)
    levels, patches = [], []
    u0 = rhs = None
    for l in range(nl):
        rp, col = M.patch_structure(l)
        lift = None
        if l == nl - 1:
            X = M.coords(l)
            lift = np.zeros((len(rp) - 1, 4))
            inflow = M.dirichlet(l, (1,))
            lift[inflow, 1] = inflow_profile(X[inflow], amp)
            walls = M.dirichlet(l, (2,))
            lift[walls, 1:] = 0.0
        val, r = M.assemble_stokes(l, len(col), visc, alpha, colours, comps, lift)
        if lift is not None:
            u0, rhs = lift, r
        lev = Level(n=len(rp) - 1, ncomp=4, rowptr=rp, col=col, val=val,
                    transfer=M.transfer(l) if l > 0 else None, hn_kind=13,
                    dirichlet=M.dirichlet(l, colours),
                    dirichlet_comps=np.asarray(comps, np.int32),
                    coords=M.coords(l) if keep_coords else None)
        levels.append(lev)
        patches.append(M.patches(l))
    h = Hierarchy(levels, 3, name or f"stokes:{inp}:{prerefine}")
    h.mesh = M
    h.patches = patches
    h.u0, h.rhs, h.amplitude = u0, rhs, amp
    h.visc, h.alpha = visc, alpha
    h.mass_in = M.face_mass(nl - 1, 1)
    h.mass_out = M.face_mass(nl - 1, 3)
    h.full_val = None
    if with_full:
        h.full_val, _ = M.assemble_stokes(nl - 1, h.fine.nnz, visc, alpha, (), ())
    # inflow data: u_x at the inflow nodes that no wall constrains to zero
    walls = np.zeros(h.fine.n, bool)
    walls[M.dirichlet(nl - 1, (2,))] = True
    inflow = M.dirichlet(nl - 1, (1,))
    h.inflow_nodes = inflow[~walls[inflow]]
    return h


def q2_hierarchy(prerefine, inp="cube.inp", drop_coarse=1, colours=(1,), name=None,
                 keep_coords=True, adaptive=0, radius=0.25, centre=(0.5, 0.5, 0.5), gup=None):
    """Poisson with the triquadratic patch element (CGDiscQ23d, Laplace<3>) on a
    refined hex mesh: one unknown per mesh node, 2x2x2-cell patches as
    elements (up to 125 couplings per row), nodal transfers between levels
    (MgInterpolatorMatrix, as for every discretisation of the reference's
    multilevel solver), Dirichlet nodes of the given colours.  With
    adaptive = 1 or 2 the finest uniform mesh is refined that many times as
    MeshAgent::refine_nodes refines the nodes inside the ball (radius,
    centre), and the multigrid levels are those Gascoigne's mesh constructor
    derives from the result (coarser levels coarsen the whole mesh where they
    can), with Q2 hanging-node constraints and the operator condensed over
    them.  With gup, the mesh hierarchy comes from that Gascoigne hierarchical
    mesh file (prerefine, inp and the ball are then unused)."""
    if gup is not None:
        M = Mesh.from_gup(gup, drop_coarse)
        adaptive = 1
    elif adaptive:
        M = Mesh.adaptive(inp, prerefine, drop_coarse, adaptive, radius, centre)
    else:
        M = Mesh(inp, 3, prerefine, drop_coarse)
    levels = []
    for l in range(M.nlevels):
        if adaptive:
            rp, col, val = M.q2_condensed(l, colours)
            he, hf = M.hn(l)
        else:
            rp, col = M.patch_structure(l)
            val = M.assemble_q2(l, len(col), colours)
            he, hf = np.zeros((0, 4), np.int32), np.zeros((0, 10), np.int32)
        levels.append(Level(n=len(rp) - 1, ncomp=1, rowptr=rp, col=col, val=val,
                            transfer=M.transfer(l) if l > 0 else None, hn_kind=23,
                            hn_edges=he, hn_faces=hf,
                            dirichlet=M.dirichlet(l, colours),
                            dirichlet_comps=np.zeros(1, np.int32),
                            coords=M.coords(l) if keep_coords else None))
    tag = f"q2:{inp}:{prerefine}" + (f"+{adaptive}" if adaptive else "")
    h = Hierarchy(levels, 3, name or tag)
    h.mesh = M
    return h
