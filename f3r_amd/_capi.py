"""ctypes binding of libf3r_amd.so (declarations: include/f3r_amd.h).

The library is built in-tree by `make -C f3r_amd` (or __graft_entry__.build()).
There is no fallback: importing this module without the built library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libf3r_amd.so")

F64, F32, F16, I32, BF16 = 0, 1, 2, 3, 4
OK = 0
ERR_ARG = -1

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `make -C f3r_amd` or __graft_entry__.build()")
lib = ctypes.CDLL(LIB_PATH)

c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [("m1", c_i32), ("m2", c_i32), ("m3", c_i32), ("m4", c_i32), ("c", c_i32),
                ("restart", c_i32), ("eps", c_dbl), ("variant", c_i32), ("range_scale", c_dbl),
                ("C", c_i32), ("sigma", c_i32), ("K", c_i32), ("use_graph", c_i32), ("flags", c_i32),
                ("grid", c_i32), ("index", c_i32), ("static_weight", c_dbl), ("depth", c_i32)]


class Level(ctypes.Structure):
    _fields_ = [("m", c_i32), ("basis", c_i32), ("matrix", c_i32)]


class Inner(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("P", c_vp), ("op", c_vp), ("m4", c_i32), ("c", c_i32),
                ("static_weight", c_dbl)]


# int (*)(void *user, f3r_vec in, f3r_vec out) and int (*)(void *user, int32 i, double e)
APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, c_vp, c_vp, c_vp)
COND_FN = ctypes.CFUNCTYPE(ctypes.c_int, c_vp, c_i32, c_dbl)


class Result(ctypes.Structure):
    _fields_ = [("is_solved", c_i32), ("res_iter", c_i32), ("res_nrm2", c_dbl), ("nrm_b", c_dbl),
                ("solve_s", c_dbl), ("precond_calls", c_i64), ("outer_cycles", c_i32)]


# (name, restype, argtypes); every entry point of include/f3r_amd.h
SIGNATURES = [
    ("f3r_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("f3r_last_error", ctypes.c_char_p, []),
    ("f3r_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    ("f3r_ctx_destroy", ctypes.c_int, [c_vp]),
    ("f3r_ctx_sync", ctypes.c_int, [c_vp]),
    ("f3r_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("f3r_vec_create", ctypes.c_int, [c_vp, ctypes.c_int, c_i64, ctypes.POINTER(c_vp)]),
    ("f3r_vec_view", ctypes.c_int, [c_vp, c_i64, c_i64, ctypes.POINTER(c_vp)]),
    ("f3r_vec_destroy", ctypes.c_int, [c_vp]),
    ("f3r_vec_upload", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_vec_download", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_vec_info", ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int), P_i64, ctypes.POINTER(c_vp)]),
    ("f3r_vec_fill", ctypes.c_int, [c_vp, c_dbl]),
    ("f3r_vec_copy", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_vec_axpy", ctypes.c_int, [c_dbl, c_vp, c_vp]),
    ("f3r_vec_scale", ctypes.c_int, [c_vp, c_dbl]),
    ("f3r_vec_dot", ctypes.c_int, [c_vp, c_vp, P_dbl]),
    ("f3r_vec_nrm2", ctypes.c_int, [c_vp, P_dbl]),
    ("f3r_vec_dot_noalloc", ctypes.c_int, [c_vp, c_vp, c_vp]),
    ("f3r_csr_scaling", ctypes.c_int, [c_i32, c_vp, c_vp, c_vp]),
    ("f3r_sell_nslots", ctypes.c_int, [c_i32, c_vp, c_i32, c_i32, c_i32, P_i64]),
    ("f3r_gen_stencil27_nnz", ctypes.c_int, [c_i32, c_i32, c_i32, P_i64]),
    ("f3r_gen_stencil27", ctypes.c_int, [c_i32, c_i32, c_i32, c_dbl, c_vp, c_vp, c_vp]),
    ("f3r_gen_convdiff7_nnz", ctypes.c_int, [c_i32, c_i32, c_i32, P_i64]),
    ("f3r_gen_convdiff7", ctypes.c_int, [c_i32, c_i32, c_i32, c_dbl, c_dbl, c_dbl, ctypes.c_uint64,
                                          c_vp, c_vp, c_vp]),
    ("f3r_gen_dominant_nnz", ctypes.c_int, [c_i32, c_i32, c_i32, c_i32, P_i64]),
    ("f3r_gen_dominant", ctypes.c_int, [c_i32, c_i32, c_i32, c_i32, c_dbl, c_vp, c_vp, c_vp]),
    ("f3r_gen_uniform", ctypes.c_int, [c_i64, c_dbl, c_dbl, ctypes.c_uint64, c_vp]),
    ("f3r_mat_create_csr", ctypes.c_int, [c_vp, c_i32, c_i32, c_i64, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    ("f3r_mat_csr_get", ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("f3r_mat_shape", ctypes.c_int, [c_vp, P_i32, P_i32]),
    ("f3r_mm_read", ctypes.c_int, [c_vp, ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    ("f3r_mm_read_vector", ctypes.c_int, [ctypes.c_char_p, c_i64, c_vp, P_i64]),
    ("f3r_mm_rhs_path", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, c_i64]),
    ("f3r_gen_uniform_mt", ctypes.c_int, [c_i64, c_dbl, c_dbl, ctypes.c_uint64, c_vp]),
    ("f3r_mat_to_sell", ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_i32, c_i32, c_i32, ctypes.POINTER(c_vp)]),
    ("f3r_sell_layout", ctypes.c_int, [c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp,
                                        c_vp]),
    ("f3r_mat_cast", ctypes.c_int, [c_vp, ctypes.c_int, ctypes.POINTER(c_vp)]),
    ("f3r_mat_destroy", ctypes.c_int, [c_vp]),
    ("f3r_mat_set_threads_per_row", ctypes.c_int, [c_vp, c_i32]),
    ("f3r_mat_info", ctypes.c_int, [c_vp, P_i32, P_i64, P_i64, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]),
    ("f3r_spmv", ctypes.c_int, [c_vp, c_vp, c_vp]),
    ("f3r_residual", ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("f3r_spmv_time", ctypes.c_int, [c_vp, c_vp, c_vp, c_i32, P_dbl]),
    ("f3r_sdainv_create", ctypes.c_int, [c_vp, c_vp, c_dbl, c_dbl, ctypes.c_int, ctypes.POINTER(c_vp)]),
    ("f3r_ilu0_create", ctypes.c_int, [c_vp, c_vp, c_i32, c_dbl, ctypes.c_int, ctypes.POINTER(c_vp)]),
    ("f3r_sdainv_create_ex", ctypes.c_int, [c_vp, c_vp, c_dbl, c_dbl, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(c_vp)]),
    ("f3r_ilu0_create_ex", ctypes.c_int, [c_vp, c_vp, c_i32, c_dbl, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(c_vp)]),
    ("f3r_prec_setup_where", ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int)]),
    ("f3r_factor_levels", ctypes.c_int, [c_vp, ctypes.c_int, c_vp, P_i32]),
    ("f3r_op_create", ctypes.c_int, [c_vp, c_i32, ctypes.c_int, APPLY_FN, c_vp, ctypes.POINTER(c_vp)]),
    ("f3r_op_destroy", ctypes.c_int, [c_vp]),
    ("f3r_nest_create", ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(Level), c_i32, ctypes.POINTER(Inner),
                                       ctypes.POINTER(Params), ctypes.POINTER(c_vp)]),
    ("f3r_solve_cond", ctypes.c_int, [c_vp, c_vp, c_vp, COND_FN, c_vp, ctypes.POINTER(Result)]),
    ("f3r_batch_create", ctypes.c_int, [c_vp, c_i32, c_vp, c_vp, ctypes.POINTER(Params), ctypes.POINTER(c_vp)]),
    ("f3r_batch_solve", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(Result)]),
    ("f3r_batch_info", ctypes.c_int, [c_vp, ctypes.POINTER(c_i64), P_i32]),
    ("f3r_batch_destroy", ctypes.c_int, [c_vp]),
    ("f3r_prec_apply", ctypes.c_int, [c_vp, c_vp, c_vp]),
    ("f3r_prec_factor_nnz", ctypes.c_int, [c_vp, ctypes.c_int, P_i64]),
    ("f3r_prec_factor_get", ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_vp, c_vp]),
    ("f3r_prec_destroy", ctypes.c_int, [c_vp]),
    ("f3r_params_default", ctypes.c_int, [ctypes.POINTER(Params)]),
    ("f3r_solver_create", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(Params), ctypes.POINTER(c_vp)]),
    ("f3r_solve", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(Result)]),
    ("f3r_solver_history", ctypes.c_int, [c_vp, P_dbl, c_i32, P_i32]),
    ("f3r_solver_level_apply", ctypes.c_int, [c_vp, c_i32, c_vp, c_vp]),
    ("f3r_solver_time_kernel", ctypes.c_int, [c_vp, c_i32, c_i32, P_dbl, P_dbl]),
    ("f3r_solver_destroy", ctypes.c_int, [c_vp]),
    ("f3r_mat_to_sell_ex", ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_i32, c_i32, c_i32, c_i32,
                                          ctypes.POINTER(c_vp)]),
    ("f3r_mat_chunks", ctypes.c_int, [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    ("f3r_mat_setup_where", ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int)]),
    ("f3r_solver_setup_where", ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int)]),
    ("f3r_mat_lds_blocks", ctypes.c_int, [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    ("f3r_sell_offset_layout", ctypes.c_int, [c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, P_i64, P_i64, P_i64,
                                              c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("f3r_solver_capture", ctypes.c_int, [c_vp, c_i32, c_i32]),
    ("f3r_solver_captured", ctypes.c_int, [c_vp, ctypes.POINTER(c_i32), c_vp]),
    ("f3r_mlp_nparams", ctypes.c_int, [P_i64]),
    ("f3r_mlp_create", ctypes.c_int, [c_vp, ctypes.c_uint64, ctypes.POINTER(c_vp)]),
    ("f3r_mlp_destroy", ctypes.c_int, [c_vp]),
    ("f3r_mlp_get", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_mlp_set", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_mlp_forward", ctypes.c_int, [c_vp, c_i64, c_vp, c_vp]),
    ("f3r_mlp_grad_zero", ctypes.c_int, [c_vp]),
    ("f3r_mlp_backward", ctypes.c_int, [c_vp, c_i64, c_vp, c_vp]),
    ("f3r_mlp_grad_get", ctypes.c_int, [c_vp, c_vp]),
    ("f3r_mlp_adam", ctypes.c_int, [c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("f3r_learn_features", ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    ("f3r_lsys_create", ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    ("f3r_lsys_destroy", ctypes.c_int, [c_vp]),
    ("f3r_lsys_set_probes", ctypes.c_int, [c_vp, c_i32, c_vp]),
    ("f3r_lsys_loss", ctypes.c_int, [c_vp, c_vp, c_i32, c_i32, ctypes.POINTER(c_dbl)]),
    ("f3r_lsys_scales", ctypes.c_int, [c_vp, c_vp, c_i32, c_vp]),
    ("f3r_sdainv_rescale", ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
]

for _name, _res, _args in SIGNATURES:
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class F3RError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = lib.f3r_strerror(status).decode()
        detail = lib.f3r_last_error().decode()
        super().__init__(f"{where}: {msg} ({status}): {detail}")


def check(status, where=""):
    if status != OK:
        raise F3RError(status, where)
    return status
