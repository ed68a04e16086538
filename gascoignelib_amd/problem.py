"""Host-side description of a multigrid problem in Gascoigne's own terms.

A `Hierarchy` holds, for every multigrid level, what Gascoigne's CPU side hands
to the GPU path: the assembled block-CSR matrix (ColumnStencil layout, int32
indices, sorted columns), the mesh-transfer maps to the next coarser level
(c2f / 2er / 4er / 8er, src/Mesh/gascoignemeshtransfer.h:35-73), the
hanging-node maps (src/Mesh/hangingindexhandler.h:94-103) and the strong
Dirichlet node lists (BoundaryIndexHandler::Verteces).

Level numbering follows the solver: level 0 is the coarsest, nlevels-1 the
finest (src/Solver/stdmultilevelsolver.cc:224-231).  The mesh agent numbers the
other way round (mesh level 0 = finest); `from_fixture` converts.
"""
from dataclasses import dataclass, field

import numpy as np

HN_KIND = {"Q12d": 12, "Q13d": 13, "Q22d": 22, "Q23d": 23}


@dataclass
class Level:
    n: int
    ncomp: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    # maps from this level to the next coarser one: (c2f, zweier, vierer, achter)
    transfer: tuple = None
    hn_kind: int = 13
    hn_edges: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    hn_faces: np.ndarray = field(default_factory=lambda: np.zeros((0, 10), np.int32))
    dirichlet: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dirichlet_comps: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    coords: np.ndarray = None

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    @property
    def size(self):
        return self.n * self.ncomp


class Hierarchy:
    def __init__(self, levels, dim, name=""):
        self.levels = list(levels)
        self.dim = dim
        self.name = name

    @property
    def nlevels(self):
        return len(self.levels)

    def level(self, l):
        return self.levels[l]

    @property
    def fine(self):
        return self.levels[-1]

    @classmethod
    def from_fixture(cls, path, name=None):
        """Hierarchy in Gascoigne's own numbering, from a golden fixture written by
        tests/golden/make_fixtures.py."""
        z = np.load(path)
        nl = int(z["nlevels"])
        dim = int(z["dim"])
        disc = str(z["disc"])
        colours = [int(c) for c in z["colours"]]
        levels = []
        for s in range(nl):  # solver level s <-> mesh level nl-1-s
            m = nl - 1 - s
            p = f"L{m}_"
            dirs = [z[p + f"bnd{c}"] for c in colours if (p + f"bnd{c}") in z.files]
            dirichlet = np.concatenate(dirs).astype(np.int32) if dirs else np.zeros(0, np.int32)
            tr = None
            if m + 1 < nl:
                tr = tuple(np.ascontiguousarray(z[p + k], dtype=np.int32)
                           for k in ("c2f", "zweier", "vierer", "achter"))
            lev = Level(
                n=int(z[p + "rowptr"].shape[0] - 1), ncomp=1,
                rowptr=np.ascontiguousarray(z[p + "rowptr"], dtype=np.int32),
                col=np.ascontiguousarray(z[p + "col"], dtype=np.int32),
                val=np.ascontiguousarray(z[p + "val"], dtype=np.float64),
                transfer=tr, hn_kind=HN_KIND[disc],
                hn_edges=np.ascontiguousarray(z[p + "hn_edge"], dtype=np.int32).reshape(-1, 4),
                hn_faces=np.ascontiguousarray(z[p + "hn_face"], dtype=np.int32).reshape(-1, 10),
                dirichlet=dirichlet, dirichlet_comps=np.zeros(1, np.int32),
                coords=np.asarray(z[p + "coords"]))
            levels.append(lev)
        h = cls(levels, dim, name or str(path))
        h.fixture = z
        return h


def lexicographic_order(coords):
    """Nodes sorted by (z, y, x): internal position k -> node.  On lattice
    meshes the neighbours of a node then sit at fixed offsets (+-1, +-row,
    +-plane), which keeps the sparse gathers of every level cache-resident."""
    c = np.asarray(coords)
    keys = [np.round(c[:, d], 12) for d in range(c.shape[1])]
    return np.lexsort(keys).astype(np.int32)


def permuted_csr(level, order, sort_columns=True):
    """The level's matrix with rows and columns renumbered (internal position
    k = node order[k]), explicit zeros of the pattern kept, blocks moved
    whole.  sort_columns=False keeps each row's entries in the caller's order:
    the matrix the ordered solver's FP64 path runs (sums in the reference's
    order); sort_columns=True: ascending internal columns, the matrices of its
    reduced-precision V-cycle."""
    n = level.n
    inv = np.empty(n, np.int64)
    inv[np.asarray(order)] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(level.rowptr))
    r_new = inv[rows]
    c_new = inv[level.col]
    if sort_columns:
        idx = np.lexsort((c_new, r_new))
    else:
        idx = np.argsort(r_new, kind="stable")
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, r_new + 1, 1)
    rowptr = np.cumsum(rowptr)
    vals = np.asarray(level.val).reshape(len(level.col), -1)  # one row per (block) entry
    return (rowptr.astype(np.int32), c_new[idx].astype(np.int32),
            np.ascontiguousarray(vals[idx].reshape(-1), dtype=np.float64))


def splitmix64_uniform(seed, n, offset=0):
    """Uniform doubles in [-1,1) from SplitMix64 at positions offset..offset+n-1:
    x_k = mix(seed + (k+1)*0x9E3779B97F4A7C15), v = 2*((x>>11)*2^-53) - 1.
    Position k of the stream is independent of how many values are drawn."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def lattice_rhs(coords, seed, ncomp=1):
    """Right-hand side f_i uniform in [-1,1), keyed by the node's lattice position
    on the finest grid so that it does not depend on the numbering."""
    c = np.asarray(coords, dtype=np.float64)
    # finest spacing: smallest non-zero coordinate difference
    h = np.inf
    for d in range(c.shape[1]):
        u = np.unique(np.round(c[:, d], 12))
        if len(u) > 1:
            h = min(h, float(np.min(np.diff(u))))
    nmax = int(round(1.0 / h)) + 1
    idx = np.rint(c / h).astype(np.int64)
    key = np.zeros(len(c), dtype=np.int64)
    for d in reversed(range(c.shape[1])):
        key = key * nmax + idx[:, d]
    vals = np.empty(len(c) * ncomp)
    for comp in range(ncomp):
        v = splitmix64_uniform(seed + comp, int(key.max()) + 1)
        vals[comp::ncomp] = v[key]
    return vals


def splitmix64_normal(seed, n):
    """Standard normal samples by Box-Muller from two SplitMix64 streams."""
    u1 = 0.5 * (splitmix64_uniform(seed, n) + 1.0)          # [0, 1)
    u2 = 0.5 * (splitmix64_uniform(seed + 0x5151, n) + 1.0)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def gaussian_random_field(points, seed, sigma=1.0, length=0.1, modes=1024):
    """g(x) = sigma sqrt(2/M) sum_m cos(w_m . x + b_m), w_m ~ N(0, I / length^2),
    b_m ~ U[0, 2 pi): random Fourier features of the squared-exponential
    covariance sigma^2 exp(-|x-y|^2 / (2 length^2)) (Rahimi & Recht, 2007).
    On a tensor lattice each mode factors per axis, cos(w.x + b) =
    Re(e^{ib} prod_d e^{i w_d x_d}), and the sum becomes one complex matrix
    product; other point sets are summed directly."""
    x = np.asarray(points, dtype=np.float64)
    d = x.shape[1]
    w = splitmix64_normal(seed, modes * d).reshape(modes, d) / length
    b = np.pi * (splitmix64_uniform(seed + 0xB0B, modes) + 1.0)
    scale = sigma * np.sqrt(2.0 / modes)
    axes = [np.unique(np.round(x[:, a], 12)) for a in range(d)]
    if d in (2, 3) and np.prod([len(u) for u in axes]) == len(x):
        idx = [np.searchsorted(axes[a], np.round(x[:, a], 12)) for a in range(d)]
        E = [np.exp(1j * np.outer(w[:, a], axes[a])) for a in range(d)]  # modes x n_a
        c = np.exp(1j * b)
        if d == 2:
            G = ((c[:, None] * E[0]).T @ E[1]).real                     # n0 x n1
            return scale * G[idx[0], idx[1]]
        n0, n1 = len(axes[0]), len(axes[1])
        T = (c[:, None, None] * E[0][:, :, None] * E[1][:, None, :]).reshape(modes, n0 * n1)
        G = (T.T @ E[2]).real                                           # (n0 n1) x n2
        return scale * G[idx[0] * n1 + idx[1], idx[2]]
    out = np.zeros(len(x))
    for p0 in range(0, len(x), 32768):  # chunks keep the temporaries small
        xs = x[p0:p0 + 32768]
        for lo in range(0, modes, 256):
            out[p0:p0 + 32768] += np.cos(xs @ w[lo:lo + 256].T + b[lo:lo + 256]).sum(axis=1)
    return scale * out


def smooth_weight(points, seed, modes=2):
    """w(x) = sum_{p,q,r <= modes} a_pqr prod_d sin(p_d pi x_d), a ~ U[-1, 1)."""
    x = np.asarray(points, dtype=np.float64)
    d = x.shape[1]
    ks = np.array(np.meshgrid(*[np.arange(1, modes + 1)] * d, indexing="ij")).reshape(d, -1).T
    a = splitmix64_uniform(seed, len(ks))
    out = np.zeros(len(x))
    for coef, k in zip(a, ks):
        out += coef * np.prod(np.sin(np.pi * k * x), axis=1)
    return out


def cell_centres(coords, cells):
    return np.asarray(coords)[np.asarray(cells)].mean(axis=1)


def lumped_mass(coords, cells):
    """int phi_i over axis-aligned Q1 box cells: each corner gets vol / corners"""
    X = np.asarray(coords)
    cells = np.asarray(cells)
    npc = cells.shape[1]
    d = X.shape[1]
    lo = X[cells].min(axis=1)
    hi = X[cells].max(axis=1)
    vol = np.prod(hi - lo, axis=1)
    m = np.zeros(len(X))
    np.add.at(m, cells.ravel(), np.repeat(vol / npc, npc))
    return m


def patches_from_cells(cells, children=8):
    """Vanka patches of a refined level: the children of one coarser cell
    (consecutive groups of `children` cells) form a patch; its nodes in
    ascending order.  Every patch of a uniformly refined hex mesh has 27 nodes
    (9 for quads)."""
    cells = np.asarray(cells)
    groups = cells.reshape(-1, children * cells.shape[1])
    out = [np.unique(g) for g in groups]
    size = len(out[0])
    if any(len(o) != size for o in out):
        raise ValueError("patches of unequal size")
    return np.ascontiguousarray(np.array(out, dtype=np.int32))


def block_operator(level, C, S):
    """A block matrix of `level`'s pattern with B = len(C): block(i, j) =
    a_ij C + [i == j] S (a test operator for block kernels and the Vanka
    smoother; Dirichlet rows stay identity blocks)."""
    B = len(C)
    rows = np.repeat(np.arange(level.n), np.diff(level.rowptr))
    vals = level.val[:, None, None] * np.asarray(C)[None]
    diag = rows == level.col
    d = np.zeros(level.n, bool)
    d[level.dirichlet] = True
    vals[diag] += np.asarray(S)[None]
    vals[diag & d[rows]] = np.eye(B)
    return np.ascontiguousarray(vals.reshape(-1))
