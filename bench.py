"""Benchmark: DOF/s to 1e-10 relative residual on 3D Q1 Poisson, 128^3 cells
(2,146,689 DOF), mixed-precision vs plain FP64 MG-GMRES on one MI355X, plus
the residual SpMV's achieved HBM bandwidth against the 8 TB/s peak.

One "step" is one full solve from u = 0, stopping at
||f - A u||_2 <= 1e-10 ||f - A u0||_2 (true FP64 residual):
  * FP64: Newton / iterative refinement whose correction is one GMRES(10) cycle
    left-preconditioned by one V(2,2) cycle of damped Jacobi (omega = 0.8), dense
    FP64 direct solve on levels below 1000 nodes -- the reference's algorithm;
  * mixed: a V(3,2) cycle (the fastest setting of each path: the FP64 path is
    fastest at its reference V(2,2)) in FP32 arithmetic with FP16-stored matrix
    values, as the preconditioner of a flexible GMRES (restart length 20,
    tolerance measured from its starting residual) whose Arnoldi vectors,
    products with A, Givens rotations and update are FP64; the mixed solve
    with the identical V(2,2) cycle is reported beside it;
  * fp64_same_stop: the FP64 V-cycle in the mixed path's flexible GMRES with
    the same cycle, restart length and stopping rule -- the like-for-like FP64
    baseline of the mixed solve (the reference-algorithm FP64 solve above stops
    at its own tolerance base, its first GMRES residual estimate).
Inputs (matrix hierarchy, right-hand side) are resident in HBM before the timed
region.

usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--cells 128]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hip_events():
    lib = C.CDLL("libamdhip64.so")
    lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipEventSynchronize.argtypes = [C.c_void_p]
    lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    lib.hipEventDestroy.argtypes = [C.c_void_p]
    return lib


def time_kernel(ctx, fn, reps):
    """average duration of fn() (one kernel launch on ctx's stream), HIP events"""
    ev = hip_events()
    a, b = C.c_void_p(), C.c_void_p()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): the event runtime is not the library's")

    ok(ev.hipEventCreate(C.byref(a)), "hipEventCreate")
    ok(ev.hipEventCreate(C.byref(b)), "hipEventCreate")
    for _ in range(3):
        fn()
    ctx.sync()
    ok(ev.hipEventRecord(a, ctx.stream), "hipEventRecord")
    for _ in range(reps):
        fn()
    ok(ev.hipEventRecord(b, ctx.stream), "hipEventRecord")
    ok(ev.hipEventSynchronize(b), "hipEventSynchronize")
    ms = C.c_float()
    ok(ev.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
    ev.hipEventDestroy(a)
    ev.hipEventDestroy(b)
    return ms.value / reps * 1e-3


def spmv_bytes(n, nnz, sv, sx):
    """bytes of r = b - A x in CSR with int32 columns: values + column indices +
    row pointers, x once, b read, r written"""
    return nnz * (sv + 4) + (n + 1) * 4 + n * sx * 3


def sell_bytes(n, layout, sv, sx):
    """bytes of r = b - A x in the device layout: stored values, the column
    bytes the slices that store columns read (16- or 32-bit; none in pattern
    slices), per-slice metadata (slot, column start, column base: 12 B), one
    16-bit row length or pattern id per row, x once, b, r"""
    return (layout["stored"] * sv + layout["column_bytes"] + layout["slices"] * 12 + n * 2
            + n * sx * 3)


# pre,post smoothing sweeps per workload, (FP64 path, mixed path): each path at
# the fastest cycle measured on MI355X (V(2,2), V(3,2), V(3,3), V(4,4), V(6,6);
# V(1,1) too for the Vanka cycles).  Jacobi is a weak smoother for the
# triquadratic element and the variable coefficient: the reference's GMRES(10)
# needs a restart with V(2,2) there (18 steps) and converges in 9 with
# V(3,2) / V(3,3).  The mixed path's flexible GMRES (restart length 20) is
# fastest with V(2,2) except on the Q1 Poisson.
SMOOTH = {
    "poisson128": ("2,2", "3,2"),
    "adjoint96": ("3,3", "2,2"),
    "stokes64": ("2,2", "2,2"),
    "q2_64": ("3,2", "2,2"),
    "q2ada": ("3,3", "2,2"),
    # ILU(0) smoother, omega 1: V(1,1) is the fastest of V(1,1), V(2,2) and
    # V(4,4) (the reference's default counts, timed beside) on both paths
    "ilu128": ("1,1", "1,1"),
}


def parse_smooth(text):
    pre, post = (int(v) for v in text.split(","))
    if not (0 <= pre <= 32 and 0 <= post <= 32):
        raise SystemExit(f"bench.py: smoothing sweeps out of range: {text}")
    return pre, post


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=128, help="cells per direction (power of 2)")
    ap.add_argument("--mixed", default="f16", choices=["f32", "f16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--mixed-krylov", default="fgmres", choices=["fgmres", "gmres"],
                    help="outer solver of the mixed solve: flexible GMRES, or left-preconditioned "
                         "GMRES inside iterative refinement")
    ap.add_argument("--mixed-linear-tol", type=float, default=None,
                    help="GMRES tolerance of each outer step of the mixed solve "
                         "(default 0.9e-10 with fgmres: its estimate is the true residual; 1e-5 with gmres)")
    ap.add_argument("--mixed-smooth", default=None,
                    help="pre,post smoothing sweeps of the mixed V-cycle (default: the fastest "
                         "measured per workload, SMOOTH); the JSON also reports the reference's 2,2")
    ap.add_argument("--mixed-vmax", type=int, default=20,
                    help="restart length of the mixed path's flexible GMRES (the FP64 path keeps "
                         "the reference's GMRES(10))")
    ap.add_argument("--fp64-smooth", default=None,
                    help="pre,post smoothing sweeps of the FP64 V-cycle (default: the fastest "
                         "measured per workload, SMOOTH)")
    ap.add_argument("--order", default="lexicographic", choices=["lexicographic", "none"],
                    help="node order inside the solver (the caller's vectors keep theirs)")
    ap.add_argument("--workload", default="poisson128",
                    choices=["poisson128", "adjoint96", "stokes64", "q2_64", "q2ada", "ilu128"],
                    help="poisson128: 3D Q1 Poisson on 128^3 cells (the headline); adjoint96: "
                         "variable-coefficient diffusion on 96^3 cells, one gradient evaluation "
                         "(forward solve + adjoint solve + dJ/dk per cell) per step; stokes64: "
                         "Stokes-LPS channel on 64^3 cells, 4x4 blocks, Vanka multigrid; q2_64: "
                         "Poisson with the triquadratic patch element on 64^3 cells (125-point "
                         "rows), forward and adjoint solves; q2ada: the same on 64^3 cells "
                         "(32^3 triquadratic elements) plus 2 adaptive levels inside a seeded "
                         "ball of radius 0.25, with hanging-node constraints (1,986,413 DOF); "
                         "ilu128: the poisson128 problem with the ILU(0) smoother (the CPU "
                         "path's default) in each level's colour order")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only launch the roofline kernel (FP64 residual SpMV, then FP16/FP32) "
                         "and a plain copy of known size, for counter runs under rocprofv3")
    args = ap.parse_args()
    d64, dmx = SMOOTH[args.workload]
    args.fp64_smooth = parse_smooth(args.fp64_smooth or d64)
    args.mixed_smooth = parse_smooth(args.mixed_smooth or dmx)
    if args.mixed_linear_tol is None:
        args.mixed_linear_tol = 0.9e-10 if args.mixed_krylov == "fgmres" else 1e-5
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described, not built)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint96":
        return bench_adjoint(args)
    if args.workload == "stokes64":
        return bench_stokes(args)
    if args.workload == "q2_64":
        return bench_q2(args)
    if args.workload == "q2ada":
        return bench_q2(args, adaptive=2)
    if args.workload == "ilu128":
        return bench_ilu(args)

    from gascoignelib_amd import hip as H
    from gascoignelib_amd.problem import lattice_rhs
    from gascoignelib_amd.structure import uniform_hierarchy

    pre = int(round(np.log2(args.cells)))
    t0 = time.perf_counter()
    hier = uniform_hierarchy("cube.inp", 3, pre, drop_coarse=1)
    t_struct = time.perf_counter() - t0
    fine = hier.fine
    ndof = fine.n
    f = lattice_rhs(fine.coords, seed=2024)
    f[fine.dirichlet] = 0.0
    print(f"# hierarchy {args.cells}^3: {ndof} DOF, nnz {fine.nnz}, {hier.nlevels} levels, "
          f"built in {t_struct:.1f}s", file=sys.stderr, flush=True)

    ctx = H.Context(0)
    if args.kernel_only:
        return kernel_only(ctx, hier, f, args)
    common = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
                  ndirect=1000, order=None if args.order == "none" else args.order)
    t_solvers = {}

    def build(name, **kw):
        t0 = time.perf_counter()
        mg = H.CudaMultiLevelSolver(ctx, hier, **kw)
        t_solvers[name] = time.perf_counter() - t0
        return mg

    mg64 = build("fp64", prec="f64", **common)
    mpre, mpost = args.mixed_smooth
    common = dict(common, iterpre=args.fp64_smooth[0], iterpost=args.fp64_smooth[1])
    mixed_common = dict(common, iterpre=mpre, iterpost=mpost)
    mgmx = build("mixed", prec=args.mixed, krylov=args.mixed_krylov, **mixed_common)
    same = (mpre, mpost) == args.fp64_smooth
    mgmx22 = None if same else build("mixed_same_cycle", prec=args.mixed,
                                     krylov=args.mixed_krylov, **common)
    # the plain FP64 path run like the mixed one: the same flexible GMRES, restart
    # length, cycle and stopping rule, FP64 V-cycle (the like-for-like baseline)
    mg64f = build("fp64_same_stop", prec="f64", krylov="fgmres", **mixed_common)
    # the FP32-matrix / FP32-vector V-cycle (FP32-stored values), run like the
    # mixed one
    mgf32 = None if args.mixed == "f32" else build("mixed_fp32_matrix", prec="f32",
                                                   krylov=args.mixed_krylov, **mixed_common)
    t_setup = sum(t_solvers.values())
    F = H.CudaVectorInterface.from_global(ctx, f)
    U64, UMX = mg64.new_vector(), mgmx.new_vector()

    def solve(mg, U, linear_tol):
        U.zero()
        rep = mg.Solve(U, F, H.CGInfo(linear_tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                       vmax=10 if (mg is mg64 or (mg is not mg64f and args.mixed_krylov == "gmres"))
                       else args.mixed_vmax,
                       norm_kind=1)
        if H.N.NL_STATUS[rep.status] != "converged":
            raise RuntimeError(f"solve did not converge: {H.N.NL_STATUS[rep.status]}")
        return rep

    def timed(mg, U, label, linear_tol):
        for _ in range(args.warmup):
            solve(mg, U, linear_tol)
        ctx.sync()
        ts = []
        rep = None
        for k in range(args.steps):
            ctx.sync()
            t = time.perf_counter()
            rep = solve(mg, U, linear_tol)
            ctx.sync()
            ts.append(time.perf_counter() - t)
            print(f"# {label} step {k}: {ts[-1] * 1e3:.1f} ms, newton {rep.newton_iterations}, "
                  f"gmres {rep.gmres_iterations}", file=sys.stderr, flush=True)
        return ts, rep

    ts64, rep64 = timed(mg64, U64, "fp64", 1e-10)
    tsmx, repmx = timed(mgmx, UMX, f"mixed-{args.mixed} V({mpre},{mpost})", args.mixed_linear_tol)
    if mgmx22 is not None:
        U22 = mgmx22.new_vector()
        ts22, rep22 = timed(mgmx22, U22, f"mixed-{args.mixed} V(2,2)", args.mixed_linear_tol)
    U64F = mg64f.new_vector()
    ts64f, rep64f = timed(mg64f, U64F, f"fp64 fgmres V({mpre},{mpost})", 0.9e-10)
    if mgf32 is not None:
        UF32 = mgf32.new_vector()
        tsf32, repf32 = timed(mgf32, UF32, f"mixed-f32 V({mpre},{mpost})", args.mixed_linear_tol)
    u64, umx = U64.copy_back(), UMX.copy_back()

    # true FP64 residuals and the mixed path's error vs the FP64 solve
    R = mg64.new_vector()
    A64 = mg64.fine_matrix
    def true_rel_res(U):
        A64.residual(R, U, F)
        if mg64.fine_mask:
            mg64.fine_mask.zero(R)
        return R.norm() / F.norm()
    rr64, rrmx, rr64f = true_rel_res(U64), true_rel_res(UMX), true_rel_res(U64F)
    err_mx = float(np.linalg.norm(umx - u64) / np.linalg.norm(u64))
    f32_line = None
    if mgf32 is not None:
        tf32 = float(np.median(tsf32))
        f32_line = {"dtype": "f64 outer / f32 V-cycle (f32-stored matrix values)",
                    "cycle": "V(%d,%d)" % (mpre, mpost), "krylov": args.mixed_krylov,
                    "linear_tol": args.mixed_linear_tol, "ms_per_solve": tf32 * 1e3,
                    "dof_per_s": ndof / tf32, "gmres": repf32.gmres_iterations,
                    "true_rel_residual": true_rel_res(UF32),
                    "rel_error_vs_fp64": float(np.linalg.norm(UF32.copy_back() - u64)
                                               / np.linalg.norm(u64)),
                    "speedup_vs_fp64": float(np.median(ts64)) / tf32,
                    "speedup_vs_fp64_same_stop": float(np.median(ts64f)) / tf32}

    # roofline of the dominant kernel: the FP64 residual SpMV on the finest
    # level, on the very matrix the FP64 solve runs (the solver's internal
    # numbering; rows keep the caller's entry order, so sums follow the
    # reference's), and the mixed V-cycle's finest-level residual on its own
    # matrix (FP32 vectors, stored values, columns in internal order)
    from gascoignelib_amd.problem import lexicographic_order
    lex = (np.arange(ndof) if args.order == "none" else lexicographic_order(fine.coords))
    Ak = mg64.level_matrix(fp64=True)
    X = H.CudaVectorInterface.from_global(ctx, u64[lex])
    Fk = H.CudaVectorInterface.from_global(ctx, f[lex])
    t_spmv = time_kernel(ctx, lambda: Ak.residual(R, X, Fk), args.spmv_reps)
    lay = Ak.layout()
    b64 = sell_bytes(ndof, lay, 8, 8)
    gbs64 = b64 / t_spmv / 1e9
    csr64 = spmv_bytes(ndof, fine.nnz, 8, 8)
    Af = mgmx.level_matrix(fp64=False)
    Xf = H.CudaVectorInterface.from_global(ctx, u64[lex], 1, "f32")
    Ff = H.CudaVectorInterface.from_global(ctx, f[lex], 1, "f32")
    Rf = H.CudaVectorInterface(ctx, ndof, 1, "f32")
    t_spmv_mx = time_kernel(ctx, lambda: Af.residual(Rf, Xf, Ff), args.spmv_reps)
    bmx = sell_bytes(ndof, Af.layout(), 2 if args.mixed == "f16" else 4, 4)
    traffic = measured_traffic(lay)

    t64, tmx, t64f = float(np.median(ts64)), float(np.median(tsmx)), float(np.median(ts64f))
    dofs_mx, dofs_64 = ndof / tmx, ndof / t64

    t_asm = assembly_times(ctx, hier, mg64, mgmx)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hier, f)

    line = {
        "metric": "DOF/s to 1e-10 rel residual (mixed vs FP64 MG-GMRES); SpMV HBM GB/s vs peak",
        "value": dofs_mx,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 outer / f32 V-cycle ({args.mixed}-stored matrix values)",
        "data": "synthetic: seeded uniform [-1,1] nodal f keyed by lattice position, Dirichlet 0",
        "config": {"workload": f"3D Q1 Poisson, unit cube, {args.cells}^3 cells",
                   "dof": ndof, "nnz": fine.nnz, "levels": hier.nlevels,
                   # kept short: the record's readers truncate long strings
                   "solver": "FP64 Newton+GMRES(10) V(2,2); mixed "
                             + (f"FGMRES({args.mixed_vmax})"
                                if args.mixed_krylov == "fgmres" else "Newton+GMRES(10)")
                             + f" V({mpre},{mpost}); Jacobi 0.8; ndirect 1000",
                   "order": args.order},
        "fp64": {"dof_per_s": dofs_64, "ms_per_solve": t64 * 1e3, "linear_tol": 1e-10,
                 "newton": rep64.newton_iterations, "gmres": rep64.gmres_iterations,
                 "true_rel_residual": rr64},
        "mixed": {"dof_per_s": dofs_mx, "ms_per_solve": tmx * 1e3,
                  "newton": repmx.newton_iterations, "gmres": repmx.gmres_iterations,
                  "true_rel_residual": rrmx, "rel_error_vs_fp64": err_mx,
                  "linear_tol": args.mixed_linear_tol, "krylov": args.mixed_krylov},
        "speedup_mixed_vs_fp64": t64 / tmx,
        "fp64_same_stop": {"krylov": f"fgmres({args.mixed_vmax})",
                           "cycle": "V(%d,%d), the mixed path's" % (mpre, mpost),
                           "linear_tol": 0.9e-10, "ms_per_solve": t64f * 1e3,
                           "dof_per_s": ndof / t64f, "gmres": rep64f.gmres_iterations,
                           "true_rel_residual": rr64f,
                           "speedup_mixed_vs_this": t64f / tmx},
        "mixed_fp32_matrix": f32_line,
        "mixed_same_cycle": None if mgmx22 is None else {
            "cycle": "V(%d,%d), the FP64 path's" % args.fp64_smooth,
            "ms_per_solve": float(np.median(ts22)) * 1e3,
            "gmres": rep22.gmres_iterations,
            "speedup_vs_fp64": t64 / float(np.median(ts22))},
        # achieved: the algorithmic bytes of a CSR residual (nnz (8 + 4) +
        # (n + 1) 4 + n (8 + 8 + 8) for the FP64 residual); the layout itself
        # moves fewer (no stored zeros, shared column patterns): its own bytes
        # and the measured HBM traffic are reported beside
        "roofline": {"bound": "hbm", "achieved": csr64 / t_spmv / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": csr64 / t_spmv / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": "FP64 residual SpMV r = b - A x, finest level",
                     "bytes_per_launch": csr64, "us_per_launch": t_spmv * 1e6,
                     "pattern_entries": fine.nnz, "stored_entries": lay["stored"],
                     "format_bytes_per_launch": b64, "format_gbs": gbs64,
                     "format_frac": gbs64 / HBM_PEAK_GBS},
        "spmv_mixed": {"gbs": bmx / t_spmv_mx / 1e9, "frac": bmx / t_spmv_mx / 1e9 / HBM_PEAK_GBS,
                       "bytes_per_launch": bmx, "us_per_launch": t_spmv_mx * 1e6},
        "setup_s": dict({"structure": t_struct, "device_upload": t_setup,
                         "solvers": t_solvers}, **t_asm),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def assembly_times(ctx, hier, mg64, mgmx, reps=3):
    """Seconds to assemble the Q1 operator (k = 1): the finest level on the
    host producer (its threads) and on the device into the finest FP64 matrix
    in the caller's numbering (gh_assemble_q1), and the device re-assembly of
    every level of the FP64 and of the mixed solver (gh_mg_assemble_q1:
    operators, reduced-precision and flexible-GMRES copies, inverse diagonals,
    dense coarse inverses).  Each the median of `reps` after one warm-up; the
    values are those the solves ran on."""
    from gascoignelib_amd import hip as H
    M = hier.mesh
    fl = M.nlevels - 1
    nnz = hier.fine.nnz
    th = []
    for _ in range(reps):
        t = time.perf_counter()
        M.assemble(fl, nnz)
        th.append(time.perf_counter() - t)
    cells = [H.CoefficientCells(ctx, hier.dim, M.coords(l), M.cells(l)) for l in range(M.nlevels)]
    A = mg64.fine_matrix
    mask = mg64.fine_mask

    def timed(fn):
        fn()
        ctx.sync()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ctx.sync()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    td = timed(lambda: cells[-1].assemble(A, mask=mask))
    t64 = timed(lambda: mg64.assemble_q1(cells))
    tmx = timed(lambda: mgmx.assemble_q1(cells))
    return {"assembly_host_fine": float(np.median(th)), "assembly_device_fine": td,
            "reassembly_device_fp64_solver": t64, "reassembly_device_mixed_solver": tmx}


def measured_traffic(layout):
    """HBM bytes per launch of the FP64 residual from the committed counter
    summary (rocprofv3 FETCH_SIZE and WRITE_SIZE runs of `bench.py --kernel-only`,
    each scaled by the factor a plain copy of known size gives), when it was
    taken on the same layout; else None."""
    path = os.path.join(ROOT, "profiles", "residual_spmv_128.traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if (t.get("stored_entries") != layout["stored"]
            or t.get("column_bytes") != layout["column_bytes"]):
        return None
    return t["bytes_per_launch"]


def kernel_only(ctx, hier, f, args):
    """the roofline kernels alone (for counter runs): the FP64 solver's and the
    mixed V-cycle's finest-level residuals on their own matrices"""
    from gascoignelib_amd import hip as H
    common = dict(mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
                  ndirect=1000, order=None if args.order == "none" else args.order)
    mg64 = H.CudaMultiLevelSolver(ctx, hier, prec="f64", **common)
    mgmx = H.CudaMultiLevelSolver(ctx, hier, prec=args.mixed, krylov=args.mixed_krylov, **common)
    fine = hier.fine
    n = fine.n
    A = mg64.level_matrix(fp64=True)
    X = H.CudaVectorInterface.from_global(ctx, f)
    F = H.CudaVectorInterface.from_global(ctx, f)
    R = H.CudaVectorInterface(ctx, n)
    for _ in range(args.spmv_reps):
        A.residual(R, X, F)
    Af = mgmx.level_matrix(fp64=False)
    Xf = H.CudaVectorInterface.from_global(ctx, f, 1, "f32")
    Ff = H.CudaVectorInterface.from_global(ctx, f, 1, "f32")
    Rf = H.CudaVectorInterface(ctx, n, 1, "f32")
    for _ in range(args.spmv_reps):
        Af.residual(Rf, Xf, Ff)
    # plain copy of 2^25 doubles (256 MiB read + 256 MiB written): counter scale
    m = 1 << 25
    C1, C2 = H.CudaVectorInterface(ctx, m), H.CudaVectorInterface(ctx, m)
    for _ in range(args.spmv_reps):
        C2.copy_from(C1)
    ctx.sync()
    lf = Af.layout()
    print(json.dumps({"kernel_only": True, "reps": args.spmv_reps,
                      "residual_bytes_f64": sell_bytes(n, A.layout(), 8, 8),
                      "stored_entries": A.layout()["stored"],
                      "column_bytes": A.layout()["column_bytes"],
                      "slices": A.layout()["slices"],
                      "residual_bytes_mixed": sell_bytes(n, lf, 2 if args.mixed == "f16" else 4, 4),
                      "mixed_stored_entries": lf["stored"], "mixed_column_bytes": lf["column_bytes"],
                      "mixed_patterned_slices": lf["patterned"],
                      "copy_bytes_each_way": m * 8}), flush=True)


def q2_kernel_only(ctx, sol, n, f, args):
    """The finest-level sweeps of the Q2 solvers alone (for counter runs):
    FP64 residual and Jacobi sweep on the FP64 solver's matrix, f16-value /
    FP32-vector residual and sweep on the mixed V-cycle's matrix, each
    args.spmv_reps times, then a plain copy of 2^25 doubles (counter scale)."""
    from gascoignelib_amd import hip as H
    out = {"kernel_only": True, "reps": args.spmv_reps, "n": n}
    for tag, prec, vp, sv in (("f64", "fp64", "f64", 8), ("mixed", "mixed", "f32", 2)):
        mg = sol[prec][0]
        A = mg.level_matrix(fp64=(tag == "f64"))
        sx = 8 if vp == "f64" else 4
        X = H.CudaVectorInterface.from_global(ctx, f, 1, vp)
        B = H.CudaVectorInterface.from_global(ctx, f, 1, vp)
        R = H.CudaVectorInterface(ctx, n, 1, vp)
        for _ in range(args.spmv_reps):
            A.residual(R, X, B)
        for _ in range(args.spmv_reps):
            A.smooth(X, B, R, 1, 0.8)
        lay = A.layout()
        rb = sell_bytes(n, lay, sv, sx)
        out[tag] = {"residual_bytes": rb, "jacobi_bytes": rb + n * (sx + sx),
                    "stored_entries": lay["stored"], "slots": lay["slots"],
                    "slices": lay["slices"], "patterned_slices": lay["patterned"],
                    "column_bytes": lay["column_bytes"], "permuted": lay["permuted"]}
    m = 1 << 25
    C1, C2 = H.CudaVectorInterface(ctx, m), H.CudaVectorInterface(ctx, m)
    for _ in range(args.spmv_reps):
        C2.copy_from(C1)
    ctx.sync()
    out["copy_bytes_each_way"] = m * 8
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Variable-coefficient diffusion -div(k grad u) = f on the 3x3x3-hex unit
    cube refined to 96^3 cells (912,673 DOF, 5 levels), k = exp(g) per cell with
    g a Gaussian random field (sigma 1, correlation length 0.1, 1024 random
    Fourier modes), J = int u w dx for a smooth seeded w.  One step = forward
    solve + adjoint solve (both to 1e-10 relative FP64 residual) + dJ/dk for
    every cell; FP64 (GMRES, the reference's algorithm on A and A^T) against
    mixed precision (f16-value V-cycle inside flexible GMRES)."""
    from gascoignelib_amd import hip as H
    from gascoignelib_amd.autograd import DiffusionProblem
    from gascoignelib_amd.problem import cell_centres, gaussian_random_field

    ctx = H.Context(0)
    t0 = time.perf_counter()
    P = DiffusionProblem(ctx, inp="cube3.inp", prerefine=5, ndirect=400)
    k = np.exp(gaussian_random_field(cell_centres(P.coords, P.cells), seed=31))
    hier = P.hierarchy(k)
    t_struct = time.perf_counter() - t0
    ndof = hier.fine.n
    print(f"# 96^3 var-k: {ndof} DOF, {P.ncells} cells, {hier.nlevels} levels, "
          f"setup {t_struct:.1f}s", file=sys.stderr, flush=True)
    common = dict(mgtype="V", iterexact=10, omega=0.8, mgomega=1.0, ndirect=400,
                  order="lexicographic")
    solvers = {}
    for tag, prec, kry, tol, sm in (("fp64", "f64", "gmres", 1e-10, args.fp64_smooth),
                                    ("mixed", args.mixed, "fgmres", args.mixed_linear_tol,
                                     args.mixed_smooth)):
        c = dict(common, iterpre=sm[0], iterpost=sm[1])
        solvers[tag] = (H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=kry, **c),
                        H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=kry, adjoint=True,
                                               **c), tol)
    U, Z = H.CudaVectorInterface(ctx, ndof), H.CudaVectorInterface(ctx, ndof)
    Gk = H.CudaVectorInterface(ctx, P.ncells)

    def step(tag):
        mgF, mgA, tol = solvers[tag]
        vmax = 10 if tag == "fp64" else args.mixed_vmax
        U.zero()
        rf = mgF.Solve(U, P.F, H.CGInfo(tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                       vmax=vmax, norm_kind=1)
        Z.zero()
        ra = mgA.Solve(Z, P.G, H.CGInfo(tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                       vmax=vmax, norm_kind=1)
        P.cellmap.dJdk(Z, U, Gk)
        for r in (rf, ra):
            if H.N.NL_STATUS[r.status] != "converged":
                raise RuntimeError(f"{tag} solve did not converge")
        return rf, ra

    res = {}
    for tag in ("fp64", "mixed"):
        for _ in range(args.warmup):
            step(tag)
        ts = []
        for i in range(args.steps):
            ctx.sync()
            t = time.perf_counter()
            rf, ra = step(tag)
            ctx.sync()
            ts.append(time.perf_counter() - t)
            print(f"# {tag} step {i}: {ts[-1] * 1e3:.1f} ms (forward gmres {rf.gmres_iterations}, "
                  f"adjoint gmres {ra.gmres_iterations})", file=sys.stderr, flush=True)
        res[tag] = dict(ms=float(np.median(ts)) * 1e3, forward_gmres=rf.gmres_iterations,
                        adjoint_gmres=ra.gmres_iterations, u=U.copy_back(), z=Z.copy_back(),
                        g=Gk.copy_back())
    J = float(P.g @ res["fp64"]["u"])
    g64, gmx = res["fp64"]["g"], res["mixed"]["g"]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    t_dj = time_kernel(ctx, lambda: P.cellmap.dJdk(Z, U, Gk), args.spmv_reps)
    bytes_dj = P.ncells * (8 * 4 + 3 * 8 + 8) + ndof * 16
    # DWR indicators of the FP64 state and adjoint (the cells of each finest
    # patch are the producer's cells 8p..8p+7, so k is already in patch order)
    patches = P.mesh.patches(P.mesh.nlevels - 1)
    Dw = H.DwrPatches(ctx, 3, P.coords, patches)
    Kc = H.CudaVectorInterface.from_global(ctx, k)
    U64 = H.CudaVectorInterface.from_global(ctx, res["fp64"]["u"])
    Z64 = H.CudaVectorInterface.from_global(ctx, res["fp64"]["z"])
    eta = Dw.estimate(U64, Z64, P.F, Kc)
    t_dwr = time_kernel(ctx, lambda: Dw.estimate(U64, Z64, P.F, Kc, eta), args.spmv_reps)
    eta_h = eta.copy_back()
    line = {
        "metric": "gradient evaluations: DOF/s of forward + adjoint solve to 1e-10 rel residual "
                  "+ dJ/dk per cell (mixed vs FP64)",
        "value": ndof / (res["mixed"]["ms"] * 1e-3), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f64 outer / f32 V-cycle ({args.mixed}-stored matrix values)",
        "data": "synthetic: k = exp(GRF) per cell, seeded f keyed by lattice position, "
                "smooth seeded w",
        "config": {"workload": "3D variable-coefficient diffusion, Q1, 96^3 cells (cube3.inp)",
                   "dof": ndof, "cells": P.ncells, "levels": hier.nlevels},
        "fp64": {"dof_per_s": ndof / (res["fp64"]["ms"] * 1e-3), "ms_per_step": res["fp64"]["ms"],
                 "forward_gmres": res["fp64"]["forward_gmres"],
                 "adjoint_gmres": res["fp64"]["adjoint_gmres"]},
        "mixed": {"dof_per_s": ndof / (res["mixed"]["ms"] * 1e-3),
                  "ms_per_step": res["mixed"]["ms"],
                  "forward_gmres": res["mixed"]["forward_gmres"],
                  "adjoint_gmres": res["mixed"]["adjoint_gmres"],
                  "rel_error_u": rel(res["mixed"]["u"], res["fp64"]["u"]),
                  "rel_error_z": rel(res["mixed"]["z"], res["fp64"]["z"]),
                  "rel_error_dJdk": rel(gmx, g64)},
        "speedup_mixed_vs_fp64": res["fp64"]["ms"] / res["mixed"]["ms"],
        "J": J, "euler_identity_residual": float(abs(k @ g64 + J) / abs(J)),
        "dJdk_kernel": {"us_per_launch": t_dj * 1e6, "bytes_per_launch": bytes_dj,
                        "gbs": bytes_dj / t_dj / 1e9},
        "dwr_indicators": {"us_per_estimate": t_dwr * 1e6, "cells": int(len(eta_h)),
                           "patches": int(len(patches)), "sum_eta": float(eta_h.sum()),
                           "sum_abs_eta": float(np.abs(eta_h).sum()),
                           "note": "f is the nodal load vector used as data: timing of the "
                                   "estimator on the FP64 state and adjoint"},
        "setup_s": t_struct,
    }
    print(json.dumps(line), flush=True)


def bench_stokes(args):
    """Stokes-LPS channel (NavierStokesLps<3> at U = 0, CGDiscQ13dLps) on the
    unit cube refined to 64^3 cells: 274,625 nodes x (p, u, v, w) = 1,098,500
    unknowns, 16,974,593 4x4 blocks on the patch structure, 6 levels (2^3 ..
    64^3 cells, the 2^3 level solved directly).  Parabolic inflow of seeded
    amplitude at x = 0, no-slip walls, do-nothing outflow.  One step = one solve
    of the Newton correction from zero to 1e-10 relative FP64 residual with
    V(2,2) cycles of the Vanka smoother (omega 0.8, 2x2x2-cell patches, 108
    unknowns each): FP64 (left-preconditioned GMRES(10) in Newton, the
    reference's algorithm) against mixed precision (FP32 matrices, FP32 patch
    factors and FP32 V-cycle inside FP64 flexible GMRES(20)).  Beside each
    solve, the gradient of the outflow flux with respect to the inflow data
    (adjoint solve on the transposed levels + one transposed product) is timed
    and checked against the inflow-face masses."""
    from gascoignelib_amd import hip as H
    from gascoignelib_amd.structure import stokes_hierarchy

    cells = 64 if args.cells == 128 else args.cells
    pre = int(round(np.log2(cells)))
    t0 = time.perf_counter()
    hier = stokes_hierarchy(pre, seed=2024, with_full=True)
    t_struct = time.perf_counter() - t0
    fine = hier.fine
    nunk = fine.n * 4
    f = hier.rhs
    print(f"# stokes {cells}^3: {fine.n} nodes, {nunk} unknowns, {fine.nnz} blocks, "
          f"{hier.nlevels} levels, built in {t_struct:.1f}s", file=sys.stderr, flush=True)
    ctx = H.Context(0)
    from gascoignelib_amd.stokes import StokesFluxProblem
    order = None if args.order == "none" else args.order
    t0 = time.perf_counter()
    # forward and adjoint solvers of each precision (the adjoint on the
    # transposed levels, its Vanka factors those of the transposed patches)
    probs = {"fp64": StokesFluxProblem(ctx, hier, prec="f64", krylov="gmres", tol=1e-10,
                                       order=order, smooth=args.fp64_smooth),
             "mixed": StokesFluxProblem(ctx, hier, prec="f32", krylov="fgmres", tol=1e-10,
                                        order=order, smooth=args.mixed_smooth,
                                        vmax=args.mixed_vmax)}
    t_setup = time.perf_counter() - t0
    F = H.CudaVectorInterface.from_global(ctx, f, 4)
    U = H.CudaVectorInterface(ctx, fine.n, 4)
    grad_want = hier.mass_in[hier.inflow_nodes]

    def step(tag):
        P = probs[tag]
        U.zero()
        rep = P.mg.Solve(U, F, H.CGInfo(P.lin, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                         vmax=P.vmax, norm_kind=1)
        if H.N.NL_STATUS[rep.status] != "converged":
            raise RuntimeError(f"{tag} Stokes solve did not converge")
        return rep

    def timed(fn, label):
        for _ in range(args.warmup):
            fn()
        ts, out = [], None
        for i in range(args.steps):
            ctx.sync()
            t = time.perf_counter()
            out = fn()
            ctx.sync()
            ts.append(time.perf_counter() - t)
            print(f"# {label} step {i}: {ts[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
        return float(np.median(ts)) * 1e3, out

    res = {}
    for tag in ("fp64", "mixed"):
        ms, rep = timed(lambda: step(tag), f"{tag} forward")
        u = U.copy_back()
        ms_adj, (grad, _, arep) = timed(lambda: probs[tag].gradient(), f"{tag} adjoint")
        res[tag] = dict(ms=ms, gmres=rep.gmres_iterations, newton=rep.newton_iterations, u=u,
                        ms_adjoint=ms_adj, adjoint_gmres=arep.gmres_iterations,
                        grad_err=float(np.abs(grad - grad_want).max() / grad_want.max()))
        print(f"# {tag}: forward {ms:.1f} ms (gmres {rep.gmres_iterations}), flux gradient "
              f"{ms_adj:.1f} ms (gmres {arep.gmres_iterations}, error vs inflow-face masses "
              f"{res[tag]['grad_err']:.2e})", file=sys.stderr, flush=True)
    # true FP64 residual of the mixed result and mass conservation of both
    A = probs["fp64"].mg.fine_matrix
    U.assign(res["mixed"]["u"])
    R = H.CudaVectorInterface.from_global(ctx, f, 4)
    A.vmult(R, U, -1.0, 1.0)
    r = R.copy_back()
    r.reshape(-1, 4)[fine.dirichlet, 1:] = 0.0
    fin = float(hier.mass_in @ hier.u0[:, 1])

    def flux(x):
        return float(hier.mass_out @ (x.reshape(-1, 4)[:, 1] + hier.u0[:, 1]))

    # dominant kernels, on the finest level in the solver's internal numbering:
    # the block residual and one Vanka solve, FP64 and FP32
    from gascoignelib_amd.problem import lexicographic_order, permuted_csr
    order = (np.arange(fine.n, dtype=np.int32) if args.order == "none"
             else lexicographic_order(fine.coords))
    inv = np.empty(fine.n, np.int64)
    inv[order] = np.arange(fine.n)
    pat = np.ascontiguousarray(inv[hier.patches[-1]], dtype=np.int32)
    xk = res["fp64"]["u"].reshape(-1, 4)[order].reshape(-1)
    fk = f.reshape(-1, 4)[order].reshape(-1)
    npat = len(pat)
    m = 27 * 4
    kern = {}
    for prec, sz in (("f64", 8), ("f32", 4)):
        # the matrices the solves run: FP64 (rows in the caller's entry order)
        # and the FP32 V-cycle's
        Ak = (probs["fp64"].mg.level_matrix(fp64=True) if prec == "f64"
              else probs["mixed"].mg.level_matrix(fp64=False))
        Xk = H.CudaVectorInterface.from_global(ctx, xk, 4, prec)
        Fk = H.CudaVectorInterface.from_global(ctx, fk, 4, prec)
        Rk = H.CudaVectorInterface(ctx, fine.n, 4, prec)
        t_res = time_kernel(ctx, lambda: Ak.residual(Rk, Xk, Fk), args.spmv_reps)
        lay = Ak.layout()
        # component rows: per kept block row 4 values + a 32-bit column; per
        # lane a row index and a 16-bit length; per slice an offset; x, b
        # read, r written
        b_res = (lay["stored"] * (4 * sz + 4) + lay["slices"] * 64 * 6
                 + (lay["slices"] + 1) * 4 + nunk * sz * 3)
        Vk = H.VankaSmoother(ctx, probs["fp64"].mg.level_matrix(fp64=True), pat, prec)
        t_vk = time_kernel(ctx, lambda: Vk.solve(Xk), max(5, args.spmv_reps // 5))
        # the factors once (lower triangle forward, upper and diagonal
        # backward), per patch row a gather index and the local solution
        # written and read back by the averaging, x read and written
        b_vk = npat * (m * m * sz + m * 4 + m * sz * 2) + nunk * sz * 2
        kern[f"block_residual_{prec}"] = {
            "us_per_launch": t_res * 1e6, "bytes_per_launch": b_res,
            "stored_block_rows": lay["stored"], "slots": lay["slots"],
            "gbs": b_res / t_res / 1e9, "frac": b_res / t_res / 1e9 / HBM_PEAK_GBS}
        kern[f"vanka_solve_{prec}"] = {
            "us_per_launch": t_vk * 1e6, "bytes_per_launch": b_vk,
            "gbs": b_vk / t_vk / 1e9, "frac": b_vk / t_vk / 1e9 / HBM_PEAK_GBS}
        del Vk, Ak
    line = {
        "metric": "unknowns/s of a Stokes-LPS solve to 1e-10 rel residual (mixed vs FP64)",
        "value": nunk / (res["mixed"]["ms"] * 1e-3), "unit": "unknowns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer / f32 Vanka V-cycle (f32 matrices and patch factors)",
        "data": "synthetic: parabolic inflow of seeded amplitude "
                f"{hier.amplitude:.6f}, visc {hier.visc}, alpha {hier.alpha}",
        "config": {"workload": f"3D Stokes Q1/Q1-LPS channel, {cells}^3 cells, 4x4 blocks, "
                               "Vanka V(2,2) on 2x2x2-cell patches",
                   "nodes": fine.n, "unknowns": nunk, "blocks": fine.nnz,
                   "levels": hier.nlevels},
        "fp64": {"unknowns_per_s": nunk / (res["fp64"]["ms"] * 1e-3),
                 "ms_per_step": res["fp64"]["ms"], "newton": res["fp64"]["newton"],
                 "gmres": res["fp64"]["gmres"],
                 "flux_error": abs(flux(res["fp64"]["u"]) - fin) / fin,
                 "flux_gradient_ms": res["fp64"]["ms_adjoint"],
                 "adjoint_gmres": res["fp64"]["adjoint_gmres"],
                 "flux_gradient_error": res["fp64"]["grad_err"]},
        "mixed": {"unknowns_per_s": nunk / (res["mixed"]["ms"] * 1e-3),
                  "ms_per_step": res["mixed"]["ms"], "newton": res["mixed"]["newton"],
                  "gmres": res["mixed"]["gmres"],
                  "flux_gradient_ms": res["mixed"]["ms_adjoint"],
                  "adjoint_gmres": res["mixed"]["adjoint_gmres"],
                  "flux_gradient_error": res["mixed"]["grad_err"],
                  "rel_error_vs_fp64": float(np.linalg.norm(res["mixed"]["u"] - res["fp64"]["u"])
                                             / np.linalg.norm(res["fp64"]["u"])),
                  "true_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(f)),
                  "flux_error": abs(flux(res["mixed"]["u"]) - fin) / fin},
        "speedup_mixed_vs_fp64": res["fp64"]["ms"] / res["mixed"]["ms"],
        "flux_gradient": "dJ/dg of the outflow flux w.r.t. the inflow u_x values: adjoint "
                         "solve on the transposed levels + (j - A^T z); error = max deviation "
                         "from the inflow-face masses that mass conservation fixes",
        "kernels": kern,
        "setup_s": {"structure": t_struct, "device": t_setup},
    }
    print(json.dumps(line), flush=True)


def bench_q2(args, adaptive=0):
    """Poisson with the triquadratic patch element (CGDiscQ23d, Laplace<3>) on
    the unit cube refined to 64^3 cells: 274,625 DOF, 16,974,593 entries (up to
    125 per row), 6 levels, nodal transfers, Dirichlet 0, seeded f.  One step =
    one solve to 1e-10 relative FP64 residual (FP64 GMRES in Newton vs f16-value
    FP32 V-cycle in FP64 flexible GMRES, Jacobi omega 0.8 with each path's
    smoothing counts, levels below 1000 nodes direct), then the adjoint solve
    for J = mean of the nodal values in the ball of radius 0.25 around a seeded
    centre.  The forward solves with the reference's V(2,2) are timed beside.

    With adaptive = 2 the 64^3 mesh (32^3 triquadratic elements) is followed by
    two refinements of the nodes inside that same ball as MeshAgent::refine_nodes
    makes them, with the multigrid levels Gascoigne derives from the result:
    1,986,413 DOF, 120.8 M entries, Q2 hanging-node constraints on the
    refinement interfaces (63,008 edge and 31,504 face nodes on the finest
    level) with the operator condensed
    over them; f and the functional are distributed onto the parents of the
    hanging nodes (HNDistribute) before the solves."""
    from gascoignelib_amd import hip as H
    from gascoignelib_amd.problem import (lattice_rhs, lexicographic_order, permuted_csr,
                                          splitmix64_uniform)
    from gascoignelib_amd.structure import q2_hierarchy

    cells = 64 if args.cells == 128 else args.cells
    centre = 0.25 + 0.5 * (0.5 * (splitmix64_uniform(4242, 3) + 1.0))
    t0 = time.perf_counter()
    hier = q2_hierarchy(int(round(np.log2(cells))), adaptive=adaptive, radius=0.25,
                        centre=tuple(centre))
    t_struct = time.perf_counter() - t0
    fine = hier.fine
    n = fine.n
    f = lattice_rhs(fine.coords, seed=2024)
    inside = np.linalg.norm(fine.coords - centre, axis=1) < 0.25
    g = inside / max(1, int(inside.sum()))
    nhn = (len(fine.hn_edges), len(fine.hn_faces))
    print(f"# q2 {cells}^3 + {adaptive} adaptive: {n} DOF, nnz {fine.nnz}, {hier.nlevels} "
          f"levels, hanging nodes {nhn}, ball of {int(inside.sum())} nodes, built in "
          f"{t_struct:.1f}s", file=sys.stderr, flush=True)
    ctx = H.Context(0)
    common = dict(mgtype="V", iterexact=10, omega=0.8, mgomega=1.0, ndirect=1000,
                  order=None if args.order == "none" else args.order)
    paths = (("fp64", "f64", "gmres", 1e-10, args.fp64_smooth),
             ("mixed", args.mixed, "fgmres", args.mixed_linear_tol, args.mixed_smooth))
    t0 = time.perf_counter()
    sol = {}
    for tag, prec, kry, tol, sm in paths:
        c = dict(common, iterpre=sm[0], iterpost=sm[1])
        sol[tag] = (H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=kry, **c),
                    H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=kry, adjoint=True,
                                           **c), tol)
    t_setup = time.perf_counter() - t0
    if args.kernel_only:
        return q2_kernel_only(ctx, sol, n, f, args)
    F = H.CudaVectorInterface.from_global(ctx, f)
    G = H.CudaVectorInterface.from_global(ctx, g)
    if sum(nhn):
        hn = H.HNStructure(ctx, 23, n, fine.hn_edges, fine.hn_faces)
        hn.distribute(F)
        hn.distribute(G)
        hn.free()
    mask = H.DirichletMask(ctx, n, 1, fine.dirichlet)
    mask.zero(F)
    mask.zero(G)
    f, g = F.copy_back(), G.copy_back()
    U, Z = H.CudaVectorInterface(ctx, n), H.CudaVectorInterface(ctx, n)

    def solve(mg, X, B, tol):
        X.zero()
        vmax = args.mixed_vmax if mg.params.krylov == 1 else 10
        rep = mg.Solve(X, B, H.CGInfo(tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50), vmax=vmax,
                       norm_kind=1)
        if H.N.NL_STATUS[rep.status] != "converged":
            raise RuntimeError("Q2 solve did not converge")
        return rep

    res = {}
    for tag in ("fp64", "mixed"):
        mgF, mgA, tol = sol[tag]
        out = {}
        for what, mg, X, B in (("forward", mgF, U, F), ("adjoint", mgA, Z, G)):
            for _ in range(args.warmup):
                solve(mg, X, B, tol)
            ts = []
            for i in range(args.steps):
                ctx.sync()
                t = time.perf_counter()
                rep = solve(mg, X, B, tol)
                ctx.sync()
                ts.append(time.perf_counter() - t)
                print(f"# {tag} {what} step {i}: {ts[-1] * 1e3:.2f} ms (gmres "
                      f"{rep.gmres_iterations})", file=sys.stderr, flush=True)
            out[what] = dict(ms=float(np.median(ts)) * 1e3, gmres=rep.gmres_iterations,
                             x=X.copy_back())
        res[tag] = out
    # the reference's V(2,2) cycle beside each path's own, forward solves
    ref_cycle = {}
    for tag, prec, kry, tol, sm in paths:
        if sm == (2, 2):
            ref_cycle[tag] = dict(ms=res[tag]["forward"]["ms"], gmres=res[tag]["forward"]["gmres"])
            continue
        mg = H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=kry,
                                    **dict(common, iterpre=2, iterpost=2))
        for _ in range(args.warmup):
            solve(mg, U, F, tol)
        ts = []
        for _ in range(args.steps):
            ctx.sync()
            t = time.perf_counter()
            rep = solve(mg, U, F, tol)
            ctx.sync()
            ts.append(time.perf_counter() - t)
        ref_cycle[tag] = dict(ms=float(np.median(ts)) * 1e3, gmres=rep.gmres_iterations)
        print(f"# {tag} forward with V(2,2): {ref_cycle[tag]['ms']:.2f} ms (gmres "
              f"{rep.gmres_iterations})", file=sys.stderr, flush=True)
        mg.free()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    # the FP64 residual SpMV on the finest level, internal numbering
    order = (np.arange(n, dtype=np.int32) if args.order == "none"
             else lexicographic_order(fine.coords))
    Ak = sol["fp64"][0].level_matrix(fp64=True)  # the matrix the FP64 solve runs
    X = H.CudaVectorInterface.from_global(ctx, res["fp64"]["forward"]["x"][order])
    Fk = H.CudaVectorInterface.from_global(ctx, f[order])
    R = H.CudaVectorInterface(ctx, n)
    t_spmv = time_kernel(ctx, lambda: Ak.residual(R, X, Fk), args.spmv_reps)
    lay = Ak.layout()
    b64 = sell_bytes(n, lay, 8, 8)
    # the mixed V-cycle's finest-level residual on its own matrix (FP32 vectors)
    Am = sol["mixed"][0].level_matrix(fp64=False)
    Xm = H.CudaVectorInterface.from_global(ctx, res["fp64"]["forward"]["x"][order], 1, "f32")
    Fm = H.CudaVectorInterface.from_global(ctx, f[order], 1, "f32")
    Rm = H.CudaVectorInterface(ctx, n, 1, "f32")
    t_spmv_mx = time_kernel(ctx, lambda: Am.residual(Rm, Xm, Fm), args.spmv_reps)
    lay_mx = Am.layout()
    bmx = sell_bytes(n, lay_mx, 2 if args.mixed == "f16" else 4, 4)
    line = {
        "metric": "DOF/s of a Q2 Poisson solve to 1e-10 rel residual (mixed vs FP64), "
                  "plus the adjoint solve",
        "value": n / (res["mixed"]["forward"]["ms"] * 1e-3), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["mixed"]["forward"]["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": f"f64 outer / f32 V-cycle ({args.mixed}-stored matrix values)",
        "data": "synthetic: seeded uniform [-1,1] nodal f keyed by lattice position, "
                "Dirichlet 0; J = mean of u over a seeded ball of radius 0.25",
        "config": {"workload": f"3D Poisson, triquadratic patch element (Q23d), {cells}^3 cells"
                               + (f" + {adaptive} adaptive levels in a ball of radius 0.25 "
                                  "around a seeded centre, Q2 hanging nodes" if adaptive
                                  else ", uniform (no adaptive levels)"),
                   "dof": n, "nnz": fine.nnz, "levels": hier.nlevels,
                   "hanging_nodes": {"edge": nhn[0], "face": nhn[1]},
                   "centre": [float(c) for c in centre]},
        "fp64": {"forward_ms": res["fp64"]["forward"]["ms"],
                 "forward_gmres": res["fp64"]["forward"]["gmres"],
                 "adjoint_ms": res["fp64"]["adjoint"]["ms"],
                 "adjoint_gmres": res["fp64"]["adjoint"]["gmres"],
                 "dof_per_s": n / (res["fp64"]["forward"]["ms"] * 1e-3)},
        "mixed": {"forward_ms": res["mixed"]["forward"]["ms"],
                  "forward_gmres": res["mixed"]["forward"]["gmres"],
                  "adjoint_ms": res["mixed"]["adjoint"]["ms"],
                  "adjoint_gmres": res["mixed"]["adjoint"]["gmres"],
                  "rel_error_u_vs_fp64": rel(res["mixed"]["forward"]["x"],
                                             res["fp64"]["forward"]["x"]),
                  "rel_error_z_vs_fp64": rel(res["mixed"]["adjoint"]["x"],
                                             res["fp64"]["adjoint"]["x"])},
        "speedup_mixed_vs_fp64": res["fp64"]["forward"]["ms"] / res["mixed"]["forward"]["ms"],
        "cycles": {"fp64": "V(%d,%d)" % args.fp64_smooth, "mixed": "V(%d,%d)" % args.mixed_smooth},
        "reference_cycle": {"cycle": "V(2,2)", "fp64_forward_ms": ref_cycle["fp64"]["ms"],
                            "fp64_gmres": ref_cycle["fp64"]["gmres"],
                            "mixed_forward_ms": ref_cycle["mixed"]["ms"],
                            "mixed_gmres": ref_cycle["mixed"]["gmres"]},
        "J": float(g @ res["fp64"]["forward"]["x"]),
        "J_adjoint": float(f @ res["fp64"]["adjoint"]["x"]),
        "residual_spmv_f64": {"us_per_launch": t_spmv * 1e6, "bytes_per_launch": b64,
                              "gbs": b64 / t_spmv / 1e9,
                              "frac": b64 / t_spmv / 1e9 / HBM_PEAK_GBS,
                              "stored_entries": lay["stored"], "pattern_entries": fine.nnz,
                              "layout": lay},
        "residual_spmv_mixed": {"us_per_launch": t_spmv_mx * 1e6, "bytes_per_launch": bmx,
                                "gbs": bmx / t_spmv_mx / 1e9,
                                "frac": bmx / t_spmv_mx / 1e9 / HBM_PEAK_GBS, "layout": lay_mx},
        "setup_s": dict({"structure": t_struct, "device": t_setup},
                        **(q2_condensed_times(ctx, hier) if adaptive else q2_assembly_times(
                            ctx, hier, {t: sol[t][0] for t in ("fp64", "mixed")}))),
    }
    print(json.dumps(line), flush=True)


def q2_assembly_times(ctx, hier, solvers=None, reps=3):
    """Seconds to assemble the finest level's Q2 operator (Laplace on the
    triquadratic patch elements, Dirichlet rows): on the host producer (its
    threads) and on the device (gh_assemble_q2) into a matrix of the same
    pattern in the caller's numbering; and the device re-assembly of every
    level of each given solver (gh_mg_assemble_q2, the same coefficients, so
    the solvers keep their values).  Each the median of `reps` after one
    warm-up.  The device values equal the host's bit for bit (checked)."""
    from gascoignelib_amd import hip as H
    M = hier.mesh
    l = M.nlevels - 1
    rp, col = M.patch_structure(l)
    d = M.dirichlet(l, (1,))

    def host():
        return M.assemble_q2(l, len(col), (1,))

    val = host()
    A = H.CudaCSRMatrixInterface(ctx, rp, col, np.where(val != 0.0, val, 1.0))
    P = H.DwrPatches(ctx, 3, M.coords(l), M.patches(l))
    mask = H.DirichletMask(ctx, len(rp) - 1, 1, d)

    def device():
        P.assemble_q2(A, mask=mask)
        ctx.sync()

    pts = [H.DwrPatches(ctx, 3, M.coords(m), M.patches(m)) for m in range(M.nlevels)]
    runs = [("assembly_q2_host_fine", host), ("assembly_q2_device_fine", device)]
    for tag, mg in (solvers or {}).items():
        runs.append((f"reassembly_q2_device_{tag}_solver",
                     lambda mg=mg: (mg.assemble_q2(pts), ctx.sync())))
    out = {}
    for name, fn in runs:
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        out[name] = float(np.median(ts))
    if not np.array_equal(A.values(), val):
        raise RuntimeError("device Q2 assembly differs from the host producer")
    return out


def q2_condensed_times(ctx, hier, reps=3):
    """Seconds to form the finest level's Q2 operator condensed over its
    hanging nodes (Dirichlet rows): the host producer's assembly and
    condensation, against the device's (gh_assemble_q2 of the unconstrained
    operator, then gh_condense_q2) into matrices of the same patterns; each
    the median of `reps` after one warm-up.  The values are checked equal."""
    from gascoignelib_amd import hip as H
    M = hier.mesh
    l = M.nlevels - 1
    rp, col = M.patch_structure(l)
    free = M.assemble_q2(l, len(col), ())
    rpc, colc, want = M.q2_condensed(l, (1,))
    e, f = (np.ascontiguousarray(a, dtype=np.int32) for a in M.hn(l))
    A = H.CudaCSRMatrixInterface(ctx, rp, col, np.where(free != 0.0, free, 1.0))
    Ac = H.CudaCSRMatrixInterface(ctx, rpc, colc, np.where(want != 0.0, want, 1.0))
    P = H.DwrPatches(ctx, 3, M.coords(l), M.patches(l))
    mask = H.DirichletMask(ctx, len(rpc) - 1, 1, M.dirichlet(l, (1,)))

    def host():
        M.q2_condensed(l, (1,))

    def device():
        P.assemble_q2(A)
        H.condense_q2(ctx, e, f, A, Ac, mask)
        ctx.sync()

    out = {}
    for name, fn in (("assembly_q2_condensed_host_fine", host),
                     ("assembly_q2_condensed_device_fine", device)):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        out[name] = float(np.median(ts))
    if not np.array_equal(Ac.values(), want):
        raise RuntimeError("device Q2 condensation differs from the host producer")
    return out


def ilu_bytes(n, nnz, sv, sx):
    """bytes of one ILU(0) solve in the sliced layout: the off-diagonal factor
    values and their 32-bit columns, the diagonal inverses, per slot a row
    index and an entry count in each triangle, and the vectors: x read, y
    written (forward), y read, the coupled y values once per sweep, y and x
    written (backward)"""
    return (nnz - n) * (sv + 4) + n * sv + n * 16 + n * sx * 7


def bench_ilu(args):
    """The headline problem (3D Q1 Poisson, 128^3 cells, 2,146,689 DOF) solved
    with the ILU(0) smoother (SparseBlockIlu, src/LinAlg/sparseblockilu.xx),
    the CPU path's default linear_smooth, omega = 1 (its default): h = b - A x;
    ILU solve of h; x += h.  Each level's ILU order is the greedy colour order
    of its matrix graph (ghs_colour_order, 8 colours: 8 forward and 8 backward
    level sweeps); the reference's default order comes from METIS, which is not
    in this image.  One step = one solve to 1e-10: FP64 (Newton + GMRES(10),
    FP64 V-cycle and factors) and mixed (f16-value residuals, FP32 factors and
    vectors in the V-cycle, FP64 flexible GMRES(20)).  Also measured: one ILU
    solve on the finest level alone (HIP events), the factorisation, and on the
    host the oracle's FP64 ILU factorisation, ILU solve and a whole FP64 solve
    (cpu_baseline)."""
    from gascoignelib_amd import hip as H
    from gascoignelib_amd.problem import lattice_rhs
    from gascoignelib_amd.structure import colour_order, uniform_hierarchy

    pre = int(round(np.log2(args.cells)))
    t0 = time.perf_counter()
    hier = uniform_hierarchy("cube.inp", 3, pre, drop_coarse=1)
    t_struct = time.perf_counter() - t0
    fine = hier.fine
    n = fine.n
    f = lattice_rhs(fine.coords, seed=2024)
    f[fine.dirichlet] = 0.0
    t0 = time.perf_counter()
    orders = [colour_order(hier.level(l).rowptr, hier.level(l).col) for l in range(hier.nlevels)]
    t_order = time.perf_counter() - t0
    io = [o for o, _ in orders]
    print(f"# ilu {args.cells}^3: {n} DOF, nnz {fine.nnz}, {hier.nlevels} levels, colours "
          f"{[c for _, c in orders]}", file=sys.stderr, flush=True)
    ctx = H.Context(0)
    common = dict(mgtype="V", iterexact=10, omega=1.0, mgomega=1.0, ndirect=1000,
                  linear_smooth="ilu", ilu_orders=io,
                  order=None if args.order == "none" else args.order)
    F = H.CudaVectorInterface.from_global(ctx, f)

    def run(prec, krylov, sm, steps, warmup, label):
        t0 = time.perf_counter()
        mg = H.CudaMultiLevelSolver(ctx, hier, prec=prec, krylov=krylov, iterpre=sm[0],
                                    iterpost=sm[1], **common)
        t_setup = time.perf_counter() - t0
        U = mg.new_vector()
        tol = 1e-10 if krylov == "gmres" else args.mixed_linear_tol
        ts, rep = [], None
        for k in range(warmup + steps):
            U.zero()
            ctx.sync()
            t = time.perf_counter()
            rep = mg.Solve(U, F, H.CGInfo(tol, 1e-30, 100), H.NLInfo(1e-10, 1e-30, 50),
                           vmax=10 if krylov == "gmres" else args.mixed_vmax, norm_kind=1)
            ctx.sync()
            if k >= warmup:
                ts.append(time.perf_counter() - t)
                print(f"# {label} step {k - warmup}: {ts[-1] * 1e3:.1f} ms, gmres "
                      f"{rep.gmres_iterations}", file=sys.stderr, flush=True)
            if H.N.NL_STATUS[rep.status] != "converged":
                raise RuntimeError(f"{label}: {H.N.NL_STATUS[rep.status]}")
        A64 = mg.fine_matrix
        R = mg.new_vector()
        A64.residual(R, U, F)
        if mg.fine_mask:
            mg.fine_mask.zero(R)
        out = {"ms_per_solve": float(np.median(ts)) * 1e3, "gmres": rep.gmres_iterations,
               "newton": rep.newton_iterations, "true_rel_residual": R.norm() / F.norm(),
               "setup_s": t_setup, "cycle": "V(%d,%d)" % sm, "u": U.copy_back()}
        del mg
        return out

    r64 = run("f64", "gmres", args.fp64_smooth, args.steps, args.warmup, "fp64")
    rmx = run(args.mixed, "fgmres", args.mixed_smooth, args.steps, args.warmup, "mixed")
    r64f = run("f64", "fgmres", args.mixed_smooth, args.steps, args.warmup, "fp64 fgmres")
    ref = {p: run(prec, kr, (4, 4), max(1, args.steps // 2), 1, f"{p} V(4,4)")
           for p, prec, kr in (("fp64", "f64", "gmres"), ("mixed", args.mixed, "fgmres"))}
    err_mx = float(np.linalg.norm(rmx["u"] - r64["u"]) / np.linalg.norm(r64["u"]))

    # the ILU(0) solve alone on the finest level, colour order
    A = H.CudaCSRMatrixInterface(ctx, fine.rowptr, fine.col, fine.val)
    ctx.sync()
    t0 = time.perf_counter()
    I = H.SparseBlockIlu(ctx, A, io[-1])
    ctx.sync()
    t_factor = time.perf_counter() - t0
    X = H.CudaVectorInterface.from_global(ctx, f)
    X32 = H.CudaVectorInterface.from_global(ctx, f, 1, "f32")
    t_ilu = time_kernel(ctx, lambda: I.solve(X), args.spmv_reps)
    t_ilu32 = time_kernel(ctx, lambda: I.solve(X32), args.spmv_reps)
    levels = I.levels()
    del I
    b64, b32 = ilu_bytes(n, fine.nnz, 8, 8), ilu_bytes(n, fine.nnz, 4, 4)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_ilu(hier, f, io, args.fp64_smooth)

    gbs = b64 / t_ilu / 1e9
    line = {
        "metric": "DOF/s of a Q1 Poisson solve to 1e-10 rel residual with the ILU(0) smoother "
                  "(mixed vs FP64 MG-GMRES)",
        "value": n / (rmx["ms_per_solve"] * 1e-3), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": rmx["ms_per_solve"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f64 outer / f32 V-cycle and ILU factors ({args.mixed}-stored matrix values)",
        "data": "synthetic: seeded uniform [-1,1] nodal f keyed by lattice position, Dirichlet 0",
        "config": {"workload": f"3D Q1 Poisson, unit cube, {args.cells}^3 cells, ILU(0) smoother "
                               "in the colour order of each level",
                   "dof": n, "nnz": fine.nnz, "levels": hier.nlevels,
                   "colours": [c for _, c in orders], "omega": 1.0},
        "fp64": {k: v for k, v in r64.items() if k != "u"},
        "mixed": dict({k: v for k, v in rmx.items() if k != "u"}, rel_error_vs_fp64=err_mx),
        "fp64_same_stop": {k: v for k, v in r64f.items() if k != "u"},
        "speedup_mixed_vs_fp64": r64["ms_per_solve"] / rmx["ms_per_solve"],
        "speedup_mixed_vs_fp64_same_stop": r64f["ms_per_solve"] / rmx["ms_per_solve"],
        "reference_cycle": {p: {k: v for k, v in r.items() if k != "u"} for p, r in ref.items()},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one FP64 ILU(0) solve on the finest level (16 level sweeps of "
                               "k_ilu_sweep, HIP events around the whole solve)",
                     "bytes_per_launch": b64, "us_per_launch": t_ilu * 1e6,
                     "levels": {"forward": levels[0], "backward": levels[1]}},
        "ilu_fp32": {"us_per_solve": t_ilu32 * 1e6, "bytes": b32,
                     "gbs": b32 / t_ilu32 / 1e9},
        "setup_s": {"structure": t_struct, "colour_orders": t_order,
                    "ilu_factor_finest_device": t_factor},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_ilu(hier, f, io, sm, reps=2):
    """The oracle on the host: the FP64 ILU(0) factorisation and one ILU solve
    of the finest level in the colour order (serial, as SparseBlockIlu), and
    whole FP64 Newton/GMRES/MG solves with the ILU smoother (reference
    threading: OpenMP over the matrix-vector products and grid transfers)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    fine = hier.fine
    t = time.perf_counter()
    I = O.Ilu(O.Matrix(fine.rowptr, fine.col, fine.val), io[-1])
    t_factor = time.perf_counter() - t
    x = f.copy()
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        I.solve(x)
        ts.append(time.perf_counter() - t)
    del I
    O.lib().or_set_parallel_vectors(0)
    t = time.perf_counter()
    om = O.MG(hier, mgtype="V", iterpre=sm[0], iterpost=sm[1], iterexact=10, omega=1.0,
              mgomega=1.0, ndirect=1000, smoother="ilu", ilu_orders=io)
    t_setup = time.perf_counter() - t
    tsol = []
    for _ in range(reps):
        u = np.zeros_like(f)
        t = time.perf_counter()
        rc, rep = om.newton(u, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 50),
                            norm_kind=1)
        tsol.append(time.perf_counter() - t)
    dt = float(np.median(tsol))
    return {"value": len(f) / dt, "unit": "DOF/s", "cores": int(O.lib().or_num_threads()),
            "kind": "port",
            "sample": f"median of {reps} full FP64 solves, ILU V({sm[0]},{sm[1]}), {len(f)} DOF, "
                      f"{dt:.1f} s each, gmres {rep.gmres_iterations}; reference threading",
            "ilu_factor_finest_s": t_factor, "ilu_solve_finest_ms": float(np.median(ts)) * 1e3,
            "mg_setup_s": t_setup, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hier, f, reps=3):
    """The oracle's FP64 Newton/GMRES/MG solve of the same problem on the host
    cores, the median of `reps` solves per variant: the reference's threading
    (OpenMP over the matrix-vector products and grid transfers, vector
    operations serial as in nvector.h) and all loops parallel."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    om = O.MG(hier, mgtype="V", iterpre=2, iterpost=2, iterexact=10, omega=0.8, mgomega=1.0,
              ndirect=1000)
    out = {}
    for name, par in (("reference_threading", 0), ("all_loops_parallel", 1)):
        O.lib().or_set_parallel_vectors(par)
        ts = []
        for _ in range(reps):
            u = np.zeros_like(f)
            t = time.perf_counter()
            rc, rep = om.newton(u, f, O.cginfo(1e-10, 1e-30, 100), O.nlinfo(1e-10, 1e-30, 50),
                                norm_kind=1)
            ts.append(time.perf_counter() - t)
        dt = float(np.median(ts))
        out[name] = {"dof_per_s": len(f) / dt, "s_per_solve": dt,
                     "solves_s": [round(t, 3) for t in ts],
                     "newton": rep.newton_iterations, "gmres": rep.gmres_iterations}
    O.lib().or_set_parallel_vectors(0)
    ref = out["reference_threading"]
    return {"value": ref["dof_per_s"], "unit": "DOF/s", "cores": int(O.lib().or_num_threads()),
            "kind": "port",
            "sample": f"median of {reps} full FP64 solves, {len(f)} DOF, "
                      f"{ref['s_per_solve']:.1f} s each; reference threading",
            "variants": out, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


if __name__ == "__main__":
    main()
