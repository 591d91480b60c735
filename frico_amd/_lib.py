"""ctypes binding of libfrico_hip.so (the C ABI declared in include/frico_gpu.h).

There is no CPU fallback: if the library is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfrico_hip.so")

FRICO_OK = 0
FRICO_ERR_HIP = -1000
FRICO_ERR_NOMEM = -1001
FRICO_ERR_STATE = -1002
FRICO_ERR_MESH = -1003
FRICO_ERR_NOCONV = -1004
FRICO_ERR_SYNC = -1005
FRICO_ERR_LIMIT = -1006
FILL_QUAD, FILL_APPROX = 0, 1
C128, C64 = 0, 1
SOLVE_C128, SOLVE_C64_IR, SOLVE_C64_SYM_IR, SOLVE_C64_NP_IR = 0, 1, 2, 3
BATCH_GRAPHS, BATCH_AUTOTUNE = 1, 2


class frico_rwg(C.Structure):
    _fields_ = [("itminus", C.c_int64), ("itplus", C.c_int64), ("loc_eminus", C.c_int32),
                ("loc_eplus", C.c_int32), ("edge_index", C.c_int64), ("has_interface", C.c_int32),
                ("interface_tag", C.c_int32), ("length", C.c_double), ("Aminus", C.c_double),
                ("Aplus", C.c_double), ("pminus", C.c_double * 3), ("pplus", C.c_double * 3)]


class frico_solve_info(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("max_rel_residual", C.c_double)]


class frico_profile(C.Structure):
    _fields_ = [("fill_ms", C.c_double), ("rhs_ms", C.c_double), ("factor_ms", C.c_double),
                ("solve_ms", C.c_double), ("refine_ms", C.c_double), ("gemm_ms", C.c_double),
                ("gemm_flops", C.c_double), ("gemm_launches", C.c_int64), ("gemm_bytes", C.c_double),
                ("tripair_ms", C.c_double),
                ("tripair_pairs", C.c_double), ("tripair_launches", C.c_int64)]


_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol of include/frico_gpu.h
SIGNATURES = {
    "frico_last_error": (C.c_char_p, []),
    "frico_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "frico_ctx_destroy": (C.c_int, [_vp]),
    "frico_ctx_synchronize": (C.c_int, [_vp]),
    "frico_ctx_set_profiling": (C.c_int, [_vp, C.c_int]),
    "frico_ctx_get_profile": (C.c_int, [_vp, C.POINTER(frico_profile)]),
    "frico_ctx_set_block": (C.c_int, [_vp, C.c_int]),
    "frico_mesh_load": (C.c_int, [_vp, _d, C.c_int64, _i64, _i32, C.c_int64, _i64, _i32, C.c_int64,
                                  C.c_int]),
    "frico_num_unknowns": (C.c_int64, [_vp]),
    "frico_num_triangles": (C.c_int64, [_vp]),
    "frico_rwg_table": (C.c_int, [_vp, C.POINTER(frico_rwg)]),
    "frico_reorient_deltagap": (C.c_int, [_vp, _i32, C.c_int, C.c_int, _d, C.POINTER(C.c_int)]),
    "frico_fill": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "frico_symmetrize": (C.c_int, [_vp]),
    "frico_matrix_get": (C.c_int, [_vp, _d, C.c_int64]),
    "frico_matrix_set": (C.c_int, [_vp, _d, C.c_int64, C.c_int64]),
    "frico_rhs_deltagap": (C.c_int, [_vp, C.c_double, _i32, C.c_int, C.c_double, C.c_double, _d]),
    "frico_rhs_planewave": (C.c_int, [_vp, C.c_double, C.c_int64, _d, _d, _d, _d]),
    "frico_rhs_set": (C.c_int, [_vp, _d, C.c_int64, C.c_int64]),
    "frico_rhs_get": (C.c_int, [_vp, _d, C.c_int64]),
    "frico_factor": (C.c_int, [_vp, C.c_int]),
    "frico_factored_form": (C.c_int, [_vp]),
    "frico_solve": (C.c_int, [_vp, C.c_double, C.c_int, C.POINTER(frico_solve_info)]),
    "frico_residual": (C.c_int, [_vp, _d]),
    "frico_solution_get": (C.c_int, [_vp, _d, C.c_int64]),
    "frico_solution_set": (C.c_int, [_vp, _d, C.c_int64, C.c_int64]),
    "frico_num_rhs": (C.c_int64, [_vp]),
    "frico_triangle_currents": (C.c_int, [_vp, C.c_int64, _d, _d]),
    "frico_fields": (C.c_int, [_vp, C.c_double, _d, C.c_int64, _i64, C.c_double, _d, _d]),
    "frico_fields_noalloc": (C.c_int, [_vp, C.c_double, _vp, C.c_int64, C.c_int64, _vp, C.c_int64, _vp,
                                       C.c_double, _vp, _vp]),
    "frico_sweep_frequencies": (C.c_int, [C.c_double, C.c_double, C.c_double, _d, C.c_int64, _i64]),
    "frico_antenna_sweep": (C.c_int, [_vp, _d, C.c_int, C.c_int, _i32, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_int, _d, _d]),
    "frico_bistatic_sweep": (C.c_int, [_vp, _d, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, _d]),
    "frico_adjoint_set_fill": (C.c_int, [_vp, C.c_int]),
    "frico_adjoint_forward": (C.c_int, [_vp, C.c_double, _i32, C.c_int, C.c_double, C.c_double, _d, _d]),
    "frico_adjoint_backward": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _d]),
    "frico_adjoint_rcs_forward": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _d, _d, C.c_double, C.c_double,
                                            C.c_double, _d]),
    "frico_adjoint_rcs_backward": (C.c_int, [_vp, C.c_double, C.c_double, _d]),
    "frico_monostatic_sweep": (C.c_int, [_vp, _d, C.c_int, C.c_int, C.c_int64, _d, _d, C.c_double,
                                         C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, _d,
                                         C.POINTER(frico_solve_info)]),
    "frico_sampling_plane_parse": (C.c_int, [C.c_char_p, _d, _d, _d, _d]),
    "frico_sampling_plane_mesh": (C.c_int, [_d, C.c_double, C.c_double, C.c_double, _d, C.c_int64, _i64, C.c_int64,
                                            _i64, _i64]),
    "frico_write_port_sweep": (C.c_int, [C.c_char_p, C.c_double, C.c_int64, _d, _d, _d]),
    "frico_write_radiation_diagram": (C.c_int, [C.c_char_p, C.c_double, _d]),
    "frico_write_bistatic_rcs": (C.c_int, [C.c_char_p, _d]),
    "frico_format_double": (C.c_int, [C.c_double, C.c_char_p, C.c_int64]),
    "frico_write_antenna_h5": (C.c_int, [C.c_char_p, C.c_double, C.c_double, C.c_int64, _d, C.c_double, C.c_double,
                                         _d, _d]),
    "frico_write_linear_system_h5": (C.c_int, [C.c_char_p, C.c_int64, _d, C.c_int64, C.c_int64, _d, C.c_int64, _d,
                                               C.c_int64]),
    "frico_dump_linear_system": (C.c_int, [C.c_void_p, C.c_char_p]),
    "frico_msh_read": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "frico_msh_free": (C.c_int, [_vp]),
    "frico_msh_sizes": (C.c_int, [_vp, _i64, _i64, _i64, _i32]),
    "frico_msh_arrays": (C.c_int, [_vp, _d, _i64, _i32, _i64, _i32]),
    "frico_msh_group": (C.c_int, [_vp, C.c_int32, C.c_char_p, C.c_int64, _i32, _i32, _i32, C.c_int32, _i32]),
    "frico_batch_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "frico_batch_destroy": (C.c_int, [_vp]),
    "frico_batch_add_mesh": (C.c_int, [_vp, _d, C.c_int64, _i64, _i32, C.c_int64, _i64, _i32, C.c_int64, C.c_int,
                                       _i32, C.c_int]),
    "frico_batch_num_meshes": (C.c_int, [_vp]),
    "frico_batch_workers_used": (C.c_int, [_vp]),
    "frico_batch_antenna": (C.c_int, [_vp, _d, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                      C.c_double, _d, C.POINTER(frico_solve_info)]),
    "frico_batch_block_sizes": (C.c_int, [_vp, _i32]),
    "frico_batch_factored_forms": (C.c_int, [_vp, _i32]),
}

_lib = None


def load():
    """Loads the in-tree library; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m frico_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
