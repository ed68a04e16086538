"""Digests of a level hierarchy (test infrastructure).

Two kinds, computed by the same code for the compiled reference's output
(make_digests.py) and for the structure producer's hierarchy
(tests/test_structure_digests.py):

* raw: SHA-256 of each array exactly as the reference writes it (its own node
  numbering), with its shape -- a record of the reference's output;
* canonical: the same objects renumbered by node coordinates (lexicographic,
  z slowest), so that two producers that agree up to node numbering give
  equal digests: the CSR pattern (rows and columns renumbered, columns
  sorted), the cells (corners renumbered, corner order kept, rows sorted),
  the Dirichlet nodes, the mesh-transfer maps (c2f indexed by canonical coarse
  node; 2er/4er/8er rows of a canonical fine node and its sorted canonical
  parents; the Q2 hanging-node constraints: edge rows (node, 3 parents) as
  they are, face rows as the set of (parent, weight) pairs, since the 3x3
  parents of a face may be listed transposed under the symmetric weights);
  the matrix values, which two assemblies produce only to rounding, through
  weighted sums compared with a tolerance.
"""
import hashlib

import numpy as np


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def raw(arr):
    a = np.ascontiguousarray(arr)
    return {"sha256": sha(a), "shape": list(a.shape), "dtype": str(a.dtype)}


def rank_of(coords):
    """rank[node] = position of the node in the coordinate order"""
    x = np.round(np.asarray(coords, dtype=np.float64), 12)
    order = np.lexsort(tuple(x[:, d] for d in range(x.shape[1])))
    rank = np.empty(len(x), np.int64)
    rank[order] = np.arange(len(x))
    return rank


def canonical_csr(rowptr, col, rank, val=None, block=1):
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    r, c = rank[rows], rank[np.asarray(col)]
    o = np.lexsort((c, r))
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    out = {"n": int(n), "nnz": int(len(col)),
           "pattern_sha256": sha(np.concatenate([rp, c[o]]).astype(np.int64))}
    if val is not None:
        v = np.asarray(val, dtype=np.float64).reshape(len(col), -1)[o]
        rr, cc = r[o].astype(np.float64), c[o].astype(np.float64)
        w = ((rr + 2.0 * cc) % 17.0)[:, None] * (1.0 + np.arange(v.shape[1]))[None, :]
        out["values"] = {"sum": float(v.sum()), "sum_sq": float((v * v).sum()),
                         "sum_abs": float(np.abs(v).sum()), "weighted": float((w * v).sum())}
    return out


def canonical_level(coords, rowptr=None, col=None, val=None, cells=None, dirichlet=None):
    rank = rank_of(coords)
    out = {"nnodes": int(len(coords))}
    if rowptr is not None:
        out["csr"] = canonical_csr(rowptr, col, rank, val)
    if cells is not None:
        c = rank[np.asarray(cells)]
        c = c[np.lexsort(tuple(c[:, k] for k in range(c.shape[1] - 1, -1, -1)))]
        out["cells_sha256"] = sha(c.astype(np.int64))
        out["ncells"] = int(len(c))
    if dirichlet is not None:
        out["dirichlet_sha256"] = sha(np.sort(rank[np.asarray(dirichlet)]).astype(np.int64))
        out["ndirichlet"] = int(len(dirichlet))
    return out, rank


def canonical_transfer(c2f, zweier, vierer, achter, rank_f, rank_c):
    """c2f: fine node of each coarse node; zweier/vierer/achter rows: fine
    node, then 2 / 4 / 8 coarse parents"""
    out = {}
    m = np.empty(len(rank_c), np.int64)
    m[rank_c] = rank_f[np.asarray(c2f)]
    out["c2f_sha256"] = sha(m)
    for name, arr in (("zweier", zweier), ("vierer", vierer), ("achter", achter)):
        a = np.asarray(arr)
        rows = np.column_stack([rank_f[a[:, 0]], np.sort(rank_c[a[:, 1:]], axis=1)])
        rows = rows[np.argsort(rows[:, 0], kind="stable")]
        out[name + "_sha256"] = sha(rows.astype(np.int64))
        out["n" + name] = int(len(a))
    return out


Q2_EDGE_W = (0.375, 0.75, -0.125)


def canonical_hn(edges, faces, rank):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 4)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 10)
    er = rank[e] if len(e) else e
    er = er[np.lexsort(tuple(er[:, k] for k in range(3, -1, -1)))] if len(er) else er
    # face rows: the hanging node, then its parents ordered by weight class
    # (corner 1/64-type weights, edge, centre) and rank
    w = np.outer(Q2_EDGE_W, Q2_EDGE_W).ravel()
    rows = []
    for r in f:
        pairs = sorted(zip(np.round(w, 12), rank[r[1:]]))
        rows.append([rank[r[0]]] + [int(p) for _, p in pairs])
    fr = np.array(sorted(rows), dtype=np.int64).reshape(-1, 10)
    return {"hn_edge_sha256": sha(er), "nhn_edge": int(len(e)),
            "hn_face_sha256": sha(fr), "nhn_face": int(len(f))}
