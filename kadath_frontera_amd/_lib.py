"""ctypes binding of libkadath_amd.so (the C-ABI declared in include/kadath_amd.h).

There is no fallback: if the shared library is missing, importing this module
raises, and every GPU entry point goes through the HIP kernels.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkadath_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python kadath_frontera_amd/build.py`")

lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)
c_dp = ctypes.POINTER(ctypes.c_double)
c_char_p = ctypes.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    "kad_ctx_create": (c_int, [ctypes.POINTER(c_vp), c_int]),
    "kad_ctx_destroy": (c_int, [c_vp]),
    "kad_ctx_stream": (c_vp, [c_vp]),
    "kad_sync": (c_int, [c_vp]),
    "kad_gemm": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64]),
    "kad_lu_factor": (c_int, [c_vp, c_int, c_int, c_vp, c_i64, c_vp, c_vp, c_int]),
    "kad_lu_solve": (c_int, [c_vp, c_int, c_int, c_vp, c_i64, c_vp, c_vp, c_int]),
    "kad_lu_solve_noalloc": (c_int, [c_vp, c_int, c_int, c_vp, c_i64, c_vp, c_vp, c_int]),
    "kad_gemv64": (c_int, [c_vp, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_int]),
    "kad_cast_f64_to_f32": (c_int, [c_vp, c_int, c_vp, c_i64, c_vp, c_i64]),
    "kad_ir_solve": (c_int, [c_vp, c_int, c_vp, c_i64, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_dbl, c_int, c_int,
                             c_int, c_ip, c_dp]),
    "kad_coef_3d": (c_int, [c_vp, c_int, c_int, c_int, c_ip, c_ip, c_int, c_int, c_vp, c_vp]),
    "kad_coef_i_3d": (c_int, [c_vp, c_int, c_int, c_int, c_ip, c_ip, c_int, c_int, c_vp, c_vp]),
    "kad_coef_1d": (c_int, [c_vp, c_int, c_int, c_i64, c_vp, c_vp]),
    "kad_coef_i_1d": (c_int, [c_vp, c_int, c_int, c_i64, c_vp, c_vp]),
    "kad_host_transform_matrix": (c_int, [c_int, c_int, c_int, c_dp, c_int]),
    "kad_host_op1d_matrix": (c_int, [c_int, c_int, c_int, c_dp, c_int, c_ip]),
    "kad_xcts_create": (c_int, [c_vp, c_int, c_int, c_int, c_dp, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int,
                                c_int, ctypes.POINTER(c_vp)]),
    "kad_xcts_destroy": (c_int, [c_vp]),
    "kad_xcts_sizes": (c_int, [c_vp, c_ip, c_ip, c_ip]),
    "kad_xcts_layout": (c_int, [c_vp, c_ip, c_int]),
    "kad_xcts_row_layout": (c_int, [c_vp, c_ip, c_int]),
    "kad_xcts_set_guess": (c_int, [c_vp, c_dp]),
    "kad_xcts_residual": (c_int, [c_vp, c_vp]),
    "kad_xcts_jacobian": (c_int, [c_vp, c_vp, c_i64]),
    "kad_xcts_update": (c_int, [c_vp, c_vp]),
    "kad_xcts_get_coef": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_xcts_set_coef": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_xcts_observables": (c_int, [c_vp, c_dp]),
    "kad_xcts_set_param": (c_int, [c_vp, c_int, c_dbl]),
    "kad_xcts_sensitivity": (c_int, [c_vp, c_dp, c_dp]),
    "kad_xcts_do_newton": (c_int, [c_vp, c_dbl, c_dp, c_int, c_dp]),
    "kad_xcts_last_solve": (c_int, [c_vp, c_ip, c_dp]),
    "kad_xcts_perturbation": (c_int, [c_vp, ctypes.c_uint64, c_dbl, c_dp]),
    "kad_kernel_timing": (c_int, [c_vp, c_int, c_int]),
    "kad_equilibrate": (c_int, [c_vp, c_int, c_vp, c_i64, c_vp, c_vp]),
    "kad_cast_equilibrated": (c_int, [c_vp, c_int, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    "kad_ir_solve_scaled": (c_int, [c_vp, c_int, c_vp, c_i64, c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl,
                                    c_int, c_int, c_int, c_ip, c_dp]),
    "kad_vec_scale": (c_int, [c_vp, c_int, c_vp, c_vp]),
    "kad_ir_set_restart": (c_int, [c_vp, c_int]),
    "kad_ir_history": (c_int, [c_vp, c_int, c_dp, c_int]),
    "kad_kernel_timing_read": (c_int, [c_vp, c_int, c_dp, c_dp, c_ip]),
    "kad_xcts_host_info": (c_int, [c_int, c_int, c_int, c_dp, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_int, c_ip,
                                   c_ip, c_int]),
    "kad_xcts_host_dd_blocks": (c_int, [c_int, c_int, c_int, c_dp, c_int, c_int, c_int, c_int, c_ip, c_int, c_ip,
                                        c_ip]),
    # systems given as equation strings
    "kad_sysdef_create": (c_int, [c_int, c_int, c_int, c_dp, c_int, c_int, c_int, ctypes.POINTER(c_vp)]),
    "kad_sysdef_destroy": (c_int, [c_vp]),
    "kad_sysdef_add_var": (c_int, [c_vp, c_char_p, c_int]),
    "kad_sysdef_add_cst": (c_int, [c_vp, c_char_p, c_dbl]),
    "kad_sysdef_add_param": (c_int, [c_vp, c_char_p, c_dbl]),
    "kad_sysdef_add_var_double": (c_int, [c_vp, c_char_p, c_dbl]),
    "kad_sysdef_add_eq_int": (c_int, [c_vp, c_int, c_int, c_char_p]),
    "kad_sysdef_add_global": (c_int, [c_vp, c_char_p, c_int, c_int, c_char_p]),
    "kad_system_globals": (c_int, [c_vp, c_dp]),
    "kad_system_sensitivity": (c_int, [c_vp, c_int, c_dp, c_dp]),
    "kad_system_ndouble": (c_int, [c_vp]),
    "kad_system_get_double": (c_int, [c_vp, c_int, c_dp]),
    "kad_sysdef_add_cst_field": (c_int, [c_vp, c_char_p, c_int, c_dp, c_i64]),
    "kad_sysdef_set_flat_metric": (c_int, [c_vp, c_char_p]),
    "kad_sysdef_add_def": (c_int, [c_vp, c_char_p]),
    "kad_sysdef_add_eq_inside": (c_int, [c_vp, c_int, c_char_p, c_int]),
    "kad_sysdef_add_eq_bc": (c_int, [c_vp, c_int, c_int, c_char_p]),
    "kad_sysdef_add_eq_matching": (c_int, [c_vp, c_int, c_int, c_char_p]),
    "kad_sysdef_space_add_eq": (c_int, [c_vp, c_char_p, c_char_p, c_char_p]),
    "kad_sysdef_space_add_bc": (c_int, [c_vp, c_int, c_char_p]),
    "kad_sysdef_set_adapted": (c_int, [c_vp, c_int]),
    "kad_sysdef_coloc": (c_int, [c_vp, c_int, c_dp, c_dp, c_dp]),
    "kad_sysdef_host_trace": (c_int, [c_vp, c_int, c_ip, c_ip, c_int]),
    "kad_system_build": (c_int, [c_vp, c_vp, c_int, c_int, ctypes.POINTER(c_vp)]),
    "kad_system_destroy": (c_int, [c_vp]),
    "kad_system_sizes": (c_int, [c_vp, c_ip, c_ip, c_ip]),
    "kad_system_layout": (c_int, [c_vp, c_ip, c_int]),
    "kad_system_row_layout": (c_int, [c_vp, c_ip, c_int]),
    "kad_system_set_field": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_system_get_field": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_system_get_coef": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_system_set_coef": (c_int, [c_vp, c_int, c_int, c_dp]),
    "kad_system_add_unknowns": (c_int, [c_vp, c_dp]),
    "kad_system_update": (c_int, [c_vp, c_vp]),
    "kad_system_perturbation": (c_int, [c_vp, ctypes.c_uint64, c_dbl, c_dp]),
    "kad_system_set_param": (c_int, [c_vp, c_int, c_dbl]),
    "kad_system_do_newton": (c_int, [c_vp, c_dbl, c_dp, c_int, c_dp]),
    "kad_system_do_newton_linesearch": (c_int, [c_vp, c_dbl, c_dp, c_int, c_dbl, c_int, c_dp, c_ip]),
    "kad_system_last_solve": (c_int, [c_vp, c_ip, c_dp]),
    "kad_system_set_option": (c_int, [c_vp, c_int, c_dbl]),
    "kad_residual": (c_int, [c_vp, c_vp]),
    "kad_check_equations": (c_int, [c_vp, c_dp, c_int]),
    "kad_jvp_fill": (c_int, [c_vp, c_int, c_int, c_vp, c_i64, c_int]),
    "kad_lu_factor_batched": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_i64, c_i64, c_vp, c_vp, c_ip]),
    "kad_lu_factor_batched_static": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_i64, c_i64, c_vp, c_ip, c_dp]),
    "kad_lu_solve_batched": (c_int, [c_vp, c_int, c_int, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_int]),
    "kad_dd_create": (c_int, [c_vp, c_int, c_int, c_ip, c_ip, c_ip, ctypes.POINTER(c_vp)]),
    "kad_dd_factor": (c_int, [c_vp, c_vp, c_i64]),
    "kad_dd_factor_static": (c_int, [c_vp, c_vp, c_i64, c_dbl]),
    "kad_dd_factor_keep": (c_int, [c_vp, c_vp, c_i64, c_dbl]),
    "kad_dd_pivoting": (c_int, [c_vp, c_ip, c_dp]),
    "kad_dd_solve": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp, c_dbl, c_int, c_ip, c_dp]),
    "kad_dd_apply": (c_int, [c_vp, c_vp]),
    "kad_dd_info": (c_int, [c_vp, c_ip, c_dp]),
    "kad_dd_destroy": (c_int, [c_vp]),
    "kad_system_dd_blocks": (c_int, [c_vp, c_ip, c_int, c_ip, c_ip]),
    "kad_system_dd_info": (c_int, [c_vp, c_ip, c_dp]),
    "kad_system_dd_pivoting": (c_int, [c_vp, c_ip, c_dp]),
    "kad_vjp": (c_int, [c_vp, c_vp, c_vp]),
    "kad_param_tangent": (c_int, [c_vp, c_int, c_vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def declared_symbols(header=os.path.join(_HERE, "..", "include", "kadath_amd.h")):
    """Function names declared in the public header."""
    import re
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|double)\s+(kad_\w+)\s*\(", txt, flags=re.M)))


class KadError(RuntimeError):
    pass


def check(rc, what):
    if rc < 0:
        raise KadError(f"{what} failed with status {rc}")
    return rc
