"""ctypes binding of libhz.so (include/hz.h).  Loading fails loudly when the
library has not been built: there is no CPU fallback."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhz.so")

HZ_OK = 0
ERRORS = {1: "HZ_ERR_ARG", 2: "HZ_ERR_HIP", 3: "HZ_ERR_STATE", 4: "HZ_ERR_NOMEM", 5: "HZ_ERR_NOT_CONVERGED",
          6: "HZ_ERR_BREAKDOWN", 7: "HZ_ERR_SINGULAR", 8: "HZ_ERR_NONFINITE"}
PREC_FP64, PREC_C64 = 0, 1
SOLVE_DEFAULT, SOLVE_STREAM, SOLVE_CHUNKED, SOLVE_STREAM_1, SOLVE_STREAM_MFMA, SOLVE_SYM, SOLVE_TWIST = 0, 1, 2, 3, 4, 5, 6
DRIVER_FGMRES, DRIVER_GMRES_REF = 0, 1
ORTHO_REF_CONJ, ORTHO_MGS, ORTHO_MGS_SEQ = 0, 1, 2
MASK_FULL, MASK_PHYS = 0, 1


class HzDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("h", C.c_double * 3),
        ("l", C.c_double * 3),
        ("pml", C.c_int32 * 3),
        ("spg", C.c_int32 * 3),
        ("omega", C.c_double),
        ("beta", C.c_double),
        ("pml_c", C.c_double),
        ("mode", C.c_int32),
        ("kz_symbol", C.c_int32),
        ("c0_rule", C.c_int32),
        ("skip_ratio", C.c_int32),
        ("c_bg", C.c_double),
        ("skip_k0", C.c_double),
    ]


class KrylovOpts(C.Structure):
    _fields_ = [
        ("driver", C.c_int32),
        ("m", C.c_int32),
        ("max_restarts", C.c_int32),
        ("ortho", C.c_int32),
        ("mask", C.c_int32),
        ("monitor", C.c_int32),
        ("tol", C.c_double),
        ("breakdown", C.c_double),
        ("hist", C.POINTER(C.c_double)),
        ("hist_cap", C.c_int32),
        ("prec", C.c_int32),
        ("batch", C.c_int32),
        ("graph", C.c_int32),
        ("autotune", C.c_int32),
        ("chunk", C.c_int32),
    ]


class TuneInfo(C.Structure):
    _fields_ = [
        ("solve_kernel", C.c_int32),
        ("group_cap", C.c_int32),
        ("fft_lines", C.c_int32),
        ("from_cache", C.c_int32),
        ("solve_ms", C.c_double),
        ("fft_ms", C.c_double),
    ]


class HodlrInfo(C.Structure):
    _fields_ = [
        ("active", C.c_int32),
        ("smallsize", C.c_int32),
        ("levels", C.c_int32),
        ("nodes", C.c_int32),
        ("leaf_max", C.c_int32),
        ("max_rank", C.c_int32),
        ("eps", C.c_double),
        ("mean_rank", C.c_double),
        ("bytes", C.c_double),
        ("dense_bytes", C.c_double),
    ]


HODLR_FREE_DENSE = 1
HODLR_TWISTED = 2
FIELDS_AUTO = -1


class KrylovStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("restarts", C.c_int32),
        ("converged", C.c_int32),
        ("breakdown", C.c_int32),
        ("rel_res", C.c_double),
        ("res", C.c_double),
        ("diff_sol", C.c_double),
        ("n_hist", C.c_int32),
        ("reserved", C.c_int32),
        ("t_solve_s", C.c_double),
    ]


# every extern "C" entry point of include/hz.h with its argument types
_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double
_pp = C.POINTER(C.c_void_p)
SIGNATURES = {
    "hz_last_error": (C.c_char_p, []),
    "hz_version": (_i, []),
    "hz_problem_create": (_i, [C.POINTER(HzDesc), C.POINTER(C.c_double), _pp]),
    "hz_problem_destroy": (_i, [_vp]),
    "hz_problem_set_stream": (_i, [_vp, _vp]),
    "hz_problem_sync": (_i, [_vp]),
    "hz_problem_stream": (_vp, [_vp]),
    "hz_problem_info": (_i, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "hz_problem_slots": (_i, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hz_precond_factor": (_i, [_vp, _i]),
    "hz_precond_factor_multi": (_i, [_i, C.POINTER(C.c_void_p), _i]),
    "hz_autotune": (_i, [_vp, _i, _i, C.POINTER(TuneInfo)]),
    "hz_fgmres_multi": (_i, [_i, C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                             C.POINTER(C.c_void_p), C.POINTER(KrylovOpts), C.POINTER(C.POINTER(KrylovStats))]),
    "hz_precond_apply": (_i, [_vp, _i, _i, _vp, _vp]),
    "hz_precond_apply_noalloc": (_i, [_vp, _i, _i, _vp, _vp]),
    "hz_precond_apply_adjoint": (_i, [_vp, _i, _i, _vp, _vp]),
    "hz_precond_compress": (_i, [_vp, _i, _d, _i, _i]),
    "hz_precond_hodlr_info": (_i, [_vp, _i, C.POINTER(HodlrInfo)]),
    "hz_precond_hodlr_ranks": (_i, [_vp, _i, _i64, C.POINTER(C.c_int32), _i64]),
    "hz_L0_apply": (_i, [_vp, _i, _vp, _vp]),
    "hz_op_apply": (_i, [_vp, _i, _vp, _vp]),
    "hz_op_apply_adjoint": (_i, [_vp, _i, _vp, _vp]),
    "hz_op_apply_preconditioned": (_i, [_vp, _i, _i, _vp, _vp]),
    "hz_residual": (_i, [_vp, _i, _vp, _vp, _vp, _i, C.POINTER(C.c_double)]),
    "hz_fgmres": (_i, [_vp, _i, _vp, _vp, C.POINTER(KrylovOpts), C.POINTER(KrylovStats)]),
    "hz_bicgstab": (_i, [_vp, _i, _i, _vp, _vp, _i, _d, _i, _i, C.POINTER(C.c_double), _i,
                         C.POINTER(KrylovStats)]),
    "hz_misfit_forward": (_i, [_vp, _i, _vp, _i, C.POINTER(C.c_int64), _vp, C.POINTER(KrylovOpts), _i,
                               C.POINTER(C.c_double)]),
    "hz_misfit_backward": (_i, [_vp, _vp]),
    "hz_misfit_set_budget": (_i, [_vp, _d]),
    "hz_misfit_fields_stored": (_i, [_vp, C.POINTER(C.c_int)]),
    "hz_fft_z": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "hz_plane_solve": (_i, [_vp, _i, _i, _i, _vp]),
    "hz_velocity_from_file": (_i, [C.c_char_p, C.POINTER(HzDesc), C.POINTER(C.c_double)]),
    "hz_velocity_file_name": (_i, [_d, C.c_char_p, _i]),
    "hz_write_solution": (_i, [C.c_char_p, C.POINTER(HzDesc), _vp]),
    "hz_read_solution": (_i, [C.c_char_p, C.POINTER(HzDesc), _vp]),
    "hz_write_convergence": (_i, [C.c_char_p, _i, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "hz_convergence_file_name": (_i, [C.POINTER(HzDesc), _d, _d, _d, C.c_char_p, _i]),
    "hz_runge_ratio": (_i, [C.c_char_p, C.c_char_p, C.c_char_p, _i, _i, _i, _d, C.POINTER(C.c_double)]),
    "hz_beta3d_norms": (_i, [C.c_char_p, C.c_char_p, _i, _d, C.POINTER(C.c_double), _i, C.POINTER(C.c_double),
                             C.POINTER(C.c_double)]),
    "hz_write_beta_norms": (_i, [C.c_char_p, _i, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hz_beta_norm_file_name": (_i, [_i, _i, C.c_char_p, _i]),
}

_lib = None


def lib():
    """Load libhz.so (built by build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhz.so is not built (run __graft_entry__.build() or "
                               "python iterative-helmholtz-solver-with-pardiso_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class HzError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = lib().hz_last_error().decode(errors="replace")
        super().__init__("%s: %s (%s)" % (where, ERRORS.get(code, code), msg))


def check(rc, where):
    if rc != HZ_OK:
        raise HzError(rc, where)
    return rc


def make_desc(d):
    o = HzDesc()
    for a in range(3):
        o.n[a] = int(d.n[a])
        o.h[a] = float(d.h[a])
        o.l[a] = float(d.l[a])
        o.pml[a] = int(d.pml[a])
        o.spg[a] = int(d.spg[a])
    o.omega = float(d.omega)
    o.beta = float(d.beta)
    o.pml_c = float(d.pml_c)
    o.mode = int(d.mode)
    o.kz_symbol = int(d.kz_symbol)
    o.c0_rule = int(d.c0_rule)
    o.c_bg = float(d.c_bg)
    o.skip_ratio = int(getattr(d, "skip_ratio", 0))
    o.skip_k0 = float(getattr(d, "skip_k0", 0.0))
    return o
