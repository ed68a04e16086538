"""Writes tests/golden/digests_<name>.json from the compiled reference at the
workloads' full sizes, where the arrays themselves are too large to keep.

Runs only where the reference sources are present: the dumpers of
oracle/ref/Makefile (outputs in oracle/_ref/) build Gascoigne's own mesh
hierarchy and assemble on it; per level this records the raw digests of
what they write (its own node numbering) and the canonical digests
(tests/golden/digests.py) that the structure producer's hierarchy must
reproduce up to numbering.

usage: python tests/golden/make_digests.py [name ...]
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import digests as D  # noqa: E402
from make_fixtures import CONFIGS, EQUATION, REF_BIN, build  # noqa: E402

# the centre of the adaptive workload's ball (seeded: 0.25 + 0.25 (u + 1),
# u = SplitMix64 uniforms of seed 4242)
ADA_CENTRE = (0.6705279178912067, 0.4294083221038335, 0.4314458695334776)

# name: (dimension, coarse mesh, prerefine, patchdepth, disc, dirichlet colours,
#        sphere refinement (radius, centre x, y, z, steps) or None)
CASES = {
    # config 2: 3D Q1 Poisson on 128^3 cells (2,146,689 DOF)
    "q1_3d_128": (3, "cube.inp", 7, 1, "Q13d", "1", None),
    # config 5: Stokes-LPS channel on 64^3 cells (274,625 nodes, 4x4 blocks)
    "stokes_3d_64": (3, "channel.inp", 6, 1, "Q13dLps", "1,2", None),
    # config 4: Q2 Poisson, 64^3 cells + 2 refine_nodes steps of the nodes in
    # the ball of radius 0.25 around ADA_CENTRE, Q2 hanging nodes
    "q2_3d_64_ada": (3, "cube.inp", 6, 1, "Q23d", "1", (0.25,) + ADA_CENTRE + (2,)),
}


def run(name):
    dim, inp, pre, pdepth, disc, cols, sphere = CASES[name]
    extra = [] if sphere is None else [repr(float(v)) for v in sphere[:4]] + [str(sphere[4])]
    colours = [int(c) for c in cols.split(",")]
    with tempfile.TemporaryDirectory() as td:
        param = os.path.join(td, "p.param")
        with open(param, "w") as f:
            f.write("//Block Mesh\n")
            f.write(f"dimension {dim}\nprerefine {pre}\npatchdepth {pdepth}\n")
            f.write(f"gridname {os.path.join(CONFIGS, inp)}\n")
            f.write(EQUATION.get(disc, ""))
        mdir, ldir = os.path.join(td, "mesh"), os.path.join(td, "la")
        os.makedirs(mdir)
        os.makedirs(ldir)
        subprocess.check_call([os.path.join(REF_BIN, "mesh_dump"), param, mdir] + extra,
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        subprocess.check_call([os.path.join(REF_BIN, "la_dump"), param, ldir, disc, cols] + extra,
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        nl = int(open(os.path.join(mdir, "meta.txt")).read().split()[1])
        pre_m = lambda l, k: np.load(os.path.join(mdir, f"L{l}_{k}.npy"))  # noqa: E731
        has = lambda d, l, k: os.path.exists(os.path.join(d, f"L{l}_{k}.npy"))  # noqa: E731
        csr = ("stokes_" if disc == "Q13dLps" else "")
        levels, ranks = [], []
        for l in range(nl):
            coords = pre_m(l, "coords")
            cells = pre_m(l, "cells")
            bnd = np.unique(np.concatenate([pre_m(l, f"bnd{c}") for c in colours]))
            rec = {"level": l, "raw": {"coords": D.raw(coords), "cells": D.raw(cells),
                                       "dirichlet": D.raw(bnd)}}
            rp = col = val = None
            if has(ldir, l, csr + "rowptr"):
                rp = np.load(os.path.join(ldir, f"L{l}_{csr}rowptr.npy"))
                col = np.load(os.path.join(ldir, f"L{l}_{csr}col.npy"))
                val = np.load(os.path.join(ldir, f"L{l}_{csr}val.npy"))
                rec["raw"].update(rowptr=D.raw(rp), col=D.raw(col), val=D.raw(val))
            can, rank = D.canonical_level(coords, rp, col, val, cells, bnd)
            if disc == "Q23d":
                he, hf = pre_m(l, "hn_edge"), pre_m(l, "hn_face")
                rec["raw"].update(hn_edge=D.raw(he), hn_face=D.raw(hf))
                can.update(D.canonical_hn(he, hf, rank))
            rec["canonical"] = can
            levels.append(rec)
            ranks.append(rank)
            del rp, col, val
        for l in range(nl - 1):  # fine l, coarse l + 1
            t = [pre_m(l, k) for k in ("c2f", "zweier", "vierer", "achter")]
            levels[l]["raw"].update({k: D.raw(a) for k, a in
                                     zip(("c2f", "zweier", "vierer", "achter"), t)})
            levels[l]["canonical"]["transfer_from_coarser"] = D.canonical_transfer(
                *t, ranks[l], ranks[l + 1])
        out = {"case": name, "dimension": dim, "mesh": inp, "prerefine": pre,
               "patchdepth": pdepth, "disc": disc, "dirichlet_colours": colours,
               "sphere": sphere,
               "nlevels": nl, "levels": levels}
        path = os.path.join(HERE, f"digests_{name}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(f"{name}: {nl} levels -> {path}")


def main(argv):
    if not os.path.isdir("/root/reference/src"):
        sys.exit("the reference sources are not present; digests cannot be regenerated here")
    build()
    for name in (argv or list(CASES)):
        run(name)


if __name__ == "__main__":
    main(sys.argv[1:])
