"""ctypes bindings of the two native libraries of the package.

  libgascoigne_hip.so  -> include/gascoigne_hip.h  (HIP solve path)
  libgh_structure.so   -> include/gh_structure.h   (host structure producer)

Loading fails loudly: there is no CPU fallback for the solve path.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "libgascoigne_hip.so")
STRUCT_LIB = os.path.join(PKG, "libgh_structure.so")

i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p

GH_OK, GH_EINVAL, GH_EHIP, GH_ENOMEM, GH_ESINGULAR = 0, -1, -2, -3, -4
GH_F64, GH_F32, GH_F16 = 0, 1, 2
PREC = {"f64": GH_F64, "f32": GH_F32, "f16": GH_F16}


class GhCGInfo(C.Structure):
    _fields_ = [("tol", C.c_double), ("globaltol", C.c_double), ("breaktol", C.c_double),
                ("miniter", C.c_int), ("maxiter", C.c_int),
                ("status", C.c_int), ("iteration", C.c_int), ("totaliter", C.c_int),
                ("firstresidual", C.c_double), ("residual", C.c_double),
                ("aimedresidual", C.c_double), ("previousresidual", C.c_double),
                ("correction", C.c_double), ("previouscorrection", C.c_double),
                ("rate", C.c_double), ("lastrate", C.c_double)]


class GhNLInfo(C.Structure):
    _fields_ = [("tol", C.c_double), ("globaltol", C.c_double), ("breaktol", C.c_double),
                ("rho", C.c_double), ("linrho", C.c_double), ("maxresincrease", C.c_double),
                ("miniter", C.c_int), ("maxiter", C.c_int), ("maxrelax", C.c_int),
                ("status", C.c_int), ("iteration", C.c_int), ("laststepbad", C.c_int),
                ("relax", C.c_int), ("totaliter", C.c_int),
                ("firstresidual", C.c_double), ("residual", C.c_double),
                ("aimedresidual", C.c_double), ("previousresidual", C.c_double),
                ("correction", C.c_double), ("previouscorrection", C.c_double),
                ("rate", C.c_double), ("lastrate", C.c_double)]


class GhMgParams(C.Structure):
    _fields_ = [("mgtype", C.c_char), ("iterpre", C.c_int), ("iterpost", C.c_int),
                ("iterexact", C.c_int), ("omega", C.c_double), ("mgomega", C.c_double),
                ("smoother", C.c_int), ("ndirect", C.c_int), ("prec_maxiter", C.c_int),
                ("prec_tol", C.c_double), ("prec_globaltol", C.c_double), ("prec", C.c_int),
                ("use_graph", C.c_int), ("krylov", C.c_int), ("adjoint", C.c_int)]


class GhSolveReport(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("gmres_iterations", C.c_int),
                ("mgsteps", C.c_int64), ("status", C.c_int),
                ("first_residual", C.c_double), ("last_residual", C.c_double),
                ("damping_steps", C.c_int), ("damping_exploded", C.c_int)]


CG_STATUS = {0: "waiting", 1: "running", 2: "converged", 3: "too many iterations",
             4: "exploded", 5: "diverging"}
NL_STATUS = {0: "waiting", 1: "running", 2: "converged", 3: "diverged",
             4: "slow_convergence", 5: "stagnation", 6: "exploded"}

P = C.POINTER
HIP_SIGNATURES = {
    "gh_last_error": (C.c_char_p, []),
    "gh_version": (C.c_char_p, []),
    "gh_ctx_create": (C.c_int, [C.c_int, P(vp)]),
    "gh_ctx_destroy": (C.c_int, [vp]),
    "gh_ctx_sync": (C.c_int, [vp]),
    "gh_ctx_stream": (vp, [vp]),
    "gh_vec_create": (C.c_int, [vp, C.c_int64, C.c_int, C.c_int, P(vp)]),
    "gh_vec_destroy": (C.c_int, [vp]),
    "gh_vec_n": (C.c_int64, [vp]),
    "gh_vec_ncomp": (C.c_int, [vp]),
    "gh_vec_prec": (C.c_int, [vp]),
    "gh_vec_data": (vp, [vp]),
    "gh_vec_upload": (C.c_int, [vp, vp, f64p]),
    "gh_vec_download": (C.c_int, [vp, vp, f64p]),
    "gh_vec_zero": (C.c_int, [vp, vp]),
    "gh_vec_copy": (C.c_int, [vp, vp, vp]),
    "gh_vec_equ": (C.c_int, [vp, vp, C.c_double, vp]),
    "gh_vec_add": (C.c_int, [vp, vp, C.c_double, vp]),
    "gh_vec_sadd": (C.c_int, [vp, vp, C.c_double, C.c_double, vp]),
    "gh_vec_scal": (C.c_int, [vp, vp, C.c_double]),
    "gh_vec_norm": (C.c_int, [vp, vp, f64p]),
    "gh_vec_dot": (C.c_int, [vp, vp, vp, f64p]),
    "gh_vec_norm_l8": (C.c_int, [vp, vp, f64p]),
    "gh_csr_create": (C.c_int, [vp, C.c_int64, C.c_int, i32p, i32p, f64p, C.c_int, P(vp)]),
    "gh_csr_create_host": (C.c_int, [vp, C.c_int64, C.c_int, i32p, i32p, f64p, P(vp)]),
    "gh_csr_destroy": (C.c_int, [vp]),
    "gh_csr_n": (C.c_int64, [vp]),
    "gh_csr_nnz": (C.c_int64, [vp]),
    "gh_csr_prec": (C.c_int, [vp]),
    "gh_csr_layout": (C.c_int, [vp, i64p]),
    "gh_spmv": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_double]),
    "gh_spmv_t": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_double]),
    "gh_residual": (C.c_int, [vp, vp, vp, vp, vp]),
    "gh_smooth": (C.c_int, [vp, vp, vp, vp, vp, C.c_int, C.c_double, C.c_int]),
    "gh_jacobi": (C.c_int, [vp, vp, vp]),
    "gh_jacobi_noalloc": (C.c_int, [vp, vp, vp]),
    "gh_transfer_create": (C.c_int, [vp, C.c_int64, i32p, C.c_int64, i32p, C.c_int64, i32p,
                                     C.c_int64, i32p, P(vp)]),
    "gh_transfer_destroy": (C.c_int, [vp]),
    "gh_transfer_nfine": (C.c_int64, [vp]),
    "gh_restrict_zero": (C.c_int, [vp, vp, vp, vp]),
    "gh_prolongate_add": (C.c_int, [vp, vp, vp, vp]),
    "gh_hn_create": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int64, i32p, C.c_int64, i32p, P(vp)]),
    "gh_hn_destroy": (C.c_int, [vp]),
    "gh_hn_average": (C.c_int, [vp, vp, vp]),
    "gh_hn_distribute": (C.c_int, [vp, vp, vp]),
    "gh_hn_zero": (C.c_int, [vp, vp, vp]),
    "gh_mask_create": (C.c_int, [vp, C.c_int64, C.c_int, C.c_int64, i32p, C.c_int, i32p, P(vp)]),
    "gh_mask_destroy": (C.c_int, [vp]),
    "gh_mask_zero": (C.c_int, [vp, vp, vp]),
    "gh_cginfo_init": (None, [P(GhCGInfo), C.c_double, C.c_double, C.c_int]),
    "gh_cginfo_reset": (None, [P(GhCGInfo)]),
    "gh_cginfo_check": (C.c_int, [P(GhCGInfo), C.c_double, C.c_double]),
    "gh_nlinfo_init": (None, [P(GhNLInfo), C.c_double, C.c_double, C.c_int]),
    "gh_nlinfo_reset": (None, [P(GhNLInfo)]),
    "gh_nlinfo_check": (C.c_int, [P(GhNLInfo), C.c_int, C.c_double, C.c_double]),
    "gh_mg_params_default": (None, [P(GhMgParams)]),
    "gh_mg_create": (C.c_int, [vp, C.c_int, P(vp), P(vp), P(vp), P(vp), P(GhMgParams), P(vp)]),
    "gh_mg_create_ordered": (C.c_int, [vp, C.c_int, P(vp), P(vp), P(vp), P(vp), P(i32p),
                                       P(GhMgParams), P(vp)]),
    "gh_mg_destroy": (C.c_int, [vp]),
    "gh_mg_nlevels": (C.c_int, [vp]),
    "gh_mg_mgsteps": (C.c_int64, [vp]),
    "gh_mg_level_matrix": (vp, [vp, C.c_int, C.c_int]),
    "gh_mg_linear": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp, P(GhCGInfo)]),
    "gh_mg_precondition": (C.c_int, [vp, vp, C.c_int, vp, vp, C.c_int]),
    "gh_gmres_solve": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, P(GhCGInfo),
                                 P(C.c_int)]),
    "gh_newton_solve": (C.c_int, [vp, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, P(GhCGInfo),
                                  P(GhNLInfo), C.c_int, P(GhSolveReport)]),
    "gh_vanka_create": (C.c_int, [vp, vp, C.c_int64, C.c_int, i32p, C.c_int, P(vp)]),
    "gh_vanka_destroy": (C.c_int, [vp]),
    "gh_vanka_solve": (C.c_int, [vp, vp, vp]),
    "gh_mg_set_vanka": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int, i32p]),
    "gh_ilu_create": (C.c_int, [vp, vp, i32p, P(vp)]),
    "gh_ilu_destroy": (C.c_int, [vp]),
    "gh_ilu_solve": (C.c_int, [vp, vp, vp]),
    "gh_ilu_levels": (C.c_int, [vp, i32p, i32p]),
    "gh_ilu_download": (C.c_int, [vp, vp, i32p, i32p, i32p, f64p]),
    "gh_mg_set_ilu": (C.c_int, [vp, C.c_int, i32p]),
    "gh_patches_create": (C.c_int, [vp, C.c_int, C.c_int64, f64p, C.c_int64, i32p, P(vp)]),
    "gh_patches_destroy": (C.c_int, [vp]),
    "gh_dwr_estimate": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "gh_cells_create": (C.c_int, [vp, C.c_int, C.c_int64, f64p, C.c_int64, i32p, P(vp)]),
    "gh_cells_destroy": (C.c_int, [vp]),
    "gh_dJdk_cells": (C.c_int, [vp, vp, vp, vp, vp]),
    "gh_assemble_q1": (C.c_int, [vp, vp, vp, vp, vp]),
    "gh_csr_download_values": (C.c_int, [vp, vp, f64p]),
    "gh_form_q1": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "gh_assemble_q2": (C.c_int, [vp, vp, vp, vp, vp]),
    "gh_form_q2": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "gh_condense_q2": (C.c_int, [vp, C.c_int64, i32p, C.c_int64, i32p, vp, vp, vp]),
    "gh_mg_assemble_q1": (C.c_int, [vp, vp, P(vp), P(vp)]),
    "gh_mg_assemble_q2": (C.c_int, [vp, vp, P(vp), P(vp)]),
}

STRUCT_SIGNATURES = {
    "ghs_last_error": (C.c_char_p, []),
    "ghs_build": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, P(vp)]),
    "ghs_destroy": (C.c_int, [vp]),
    "ghs_nlevels": (C.c_int, [vp]),
    "ghs_level_sizes": (C.c_int, [vp, C.c_int, i64p, i64p, i64p]),
    "ghs_level_coords": (C.c_int, [vp, C.c_int, f64p]),
    "ghs_level_cells": (C.c_int, [vp, C.c_int, i32p]),
    "ghs_level_structure": (C.c_int, [vp, C.c_int, i32p, i32p]),
    "ghs_level_dirichlet_count": (C.c_int, [vp, C.c_int, C.c_int, i32p, i64p]),
    "ghs_level_dirichlet": (C.c_int, [vp, C.c_int, C.c_int, i32p, i32p]),
    "ghs_level_assemble": (C.c_int, [vp, C.c_int, f64p, C.c_int, i32p, f64p]),
    "ghs_level_transfer_sizes": (C.c_int, [vp, C.c_int, i64p, i64p, i64p, i64p]),
    "ghs_level_transfer": (C.c_int, [vp, C.c_int, i32p, i32p, i32p, i32p]),
    "ghs_structure_from_cells": (C.c_int, [C.c_int64, C.c_int64, C.c_int, i32p, i32p, i32p]),
    "ghs_colour_order": (C.c_int, [C.c_int64, i32p, i32p, i32p, i32p]),
    "ghs_build_gup": (C.c_int, [C.c_char_p, C.c_int, P(vp)]),
    "ghs_assemble_on": (C.c_int, [C.c_int, C.c_int64, f64p, C.c_int64, i32p, f64p, i32p, i32p,
                                  C.c_int64, i32p, f64p]),
    "ghs_level_patch_sizes": (C.c_int, [vp, C.c_int, i64p, i64p]),
    "ghs_level_patches": (C.c_int, [vp, C.c_int, i32p]),
    "ghs_level_patch_structure": (C.c_int, [vp, C.c_int, i32p, i32p]),
    "ghs_level_assemble_stokes": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_int, i32p,
                                            C.c_int, f64p, f64p, f64p]),
    "ghs_level_face_mass": (C.c_int, [vp, C.c_int, C.c_int, f64p]),
    "ghs_level_assemble_q2": (C.c_int, [vp, C.c_int, C.c_int, i32p, f64p]),
    "ghs_build_adaptive": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, f64p,
                                     P(vp)]),
    "ghs_level_hn_sizes": (C.c_int, [vp, C.c_int, i64p, i64p]),
    "ghs_level_hn": (C.c_int, [vp, C.c_int, i32p, i32p]),
    "ghs_level_q2_condensed_nnz": (C.c_int, [vp, C.c_int, C.c_int, i32p, i64p]),
    "ghs_level_q2_condensed": (C.c_int, [vp, C.c_int, i32p, i32p, f64p]),
}

_libs = {}


def _load(path, sigs, what):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{what} is not built: {path} is missing "
                           "(run `python -m gascoignelib_amd.build`)")
    lib = C.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[path] = lib
    return lib


def hip():
    return _load(HIP_LIB, HIP_SIGNATURES, "the HIP solve library")


def structure():
    return _load(STRUCT_LIB, STRUCT_SIGNATURES, "the structure producer")


class GhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gascoigne_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != GH_OK:
        raise GhError(rc, hip().gh_last_error().decode(errors="replace"))
    return rc


def check_s(rc):
    if rc != 0:
        raise RuntimeError("gh_structure: " + structure().ghs_last_error().decode(errors="replace"))
    return rc


def ptr(a, t=f64p):
    return a.ctypes.data_as(t)
