"""ctypes binding of libfemd_hip.so (include/femd_hip.h).

The library is built in-tree by ``python femd_amd/build.py`` (or ``__graft_entry__.build()``);
importing the package without it raises ImportError: there is no CPU fallback.
Status codes become the exceptions the reference's nanobind layer raises: FEMD_EINVAL ->
ValueError, FEMD_EZEROPIVOT -> ValueError for a zero diagonal and RuntimeError for a zero
ILU(0) pivot (precond.hpp:68-78, 181), FEMD_EPATTERN -> RuntimeError (std::logic_error),
FEMD_EDEVICE -> RuntimeError, FEMD_ENOMEM -> MemoryError.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfemd_hip.so"

OK, EINVAL, EPATTERN, EZEROPIVOT, ENOTSPD, EDEVICE, ENOMEM = range(7)
PC_NONE, PC_JACOBI, PC_ILU0, PC_SSOR = range(4)
ORDER_NATURAL, ORDER_RCM = 0, 1
FP64, FP32 = 0, 1

_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_i64 = C.c_int64
_sz = C.c_size_t
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Field(C.Structure):
    """femd_field: a known field an INPUT instruction reads (space, host coefficients, derivative code)."""
    _fields_ = [("space", C.c_void_p), ("coefficients", C.POINTER(_d)), ("code", _i)]


class Coefficient(C.Structure):
    """femd_coefficient: a register program (forms/integrand.hpp) or host values."""
    _fields_ = [("ninstr", _i), ("code", C.POINTER(_i)), ("nreg", _i), ("nconst", _i),
                ("consts", C.POINTER(_d)), ("values", C.POINTER(_d)), ("ninputs", _i), ("inputs", C.POINTER(Field))]


class SolveInfo(C.Structure):
    _fields_ = [("converged", _i), ("iterations", _i), ("breakdown", _i), ("residual", _d), ("precond_residual", _d)]


_SIGS = {
    "femd_last_error": (C.c_char_p, []),
    "femd_init": (_i, [_i]),
    "femd_dev_alloc": (_i, [_sz, C.POINTER(_p)]),
    "femd_dev_free": (_i, [_p]),
    "femd_memcpy_htod": (_i, [_p, _p, _sz]),
    "femd_memcpy_dtoh": (_i, [_p, _p, _sz]),
    "femd_stream_sync": (_i, [_p]),
    "femd_vec_copy_dev": (_i, [_i64, _p, _p, _p]),
    "femd_mesh_rectangle": (_i, [_d, _d, _d, _d, _i, _i, _i, _ip, C.POINTER(_p)]),
    "femd_mesh_from_arrays": (_i, [_i, _dp, _i, _ip, _i, _ip, _p, _i, C.POINTER(_p)]),
    "femd_quad_mesh_rectangle": (_i, [_d, _d, _d, _d, _i, _i, _ip, C.POINTER(_p)]),
    "femd_quad_mesh_from_arrays": (_i, [_i, _dp, _i, _ip, _i, _ip, _p, _i, C.POINTER(_p)]),
    "femd_mesh_cell_vertices": (_i, [_p, C.POINTER(_i)]),
    "femd_mesh_sizes": (_i, [_p, _ip]),
    "femd_mesh_arrays": (_i, [_p, _dp, _ip, _ip, _ip, _ip]),
    "femd_mesh_free": (None, [_p]),
    "femd_space_create": (_i, [_p, _i, _i, _ip, _i, _i, C.POINTER(_p)]),
    "femd_space_sizes": (_i, [_p, _ip]),
    "femd_space_numbering": (_i, [_p, _ip, _ip, _ip, _dp]),
    "femd_space_free": (None, [_p]),
    "femd_pattern": (_i, [_p, C.POINTER(_i64), _p, _p]),
    "femd_quadrature": (_i, [_i, C.POINTER(_i), _dp, _dp, _dp]),
    "femd_quadrature_square": (_i, [_i, C.POINTER(_i), _dp, _dp, _dp]),
    "femd_program_compact": (_i, [C.POINTER(Coefficient), _ip, C.POINTER(_i)]),
    "femd_csr_create_space": (_i, [_p, C.POINTER(_p)]),
    "femd_csr_create": (_i, [_i, _i, _ip, _ip, _dp, C.POINTER(_p)]),
    "femd_csr_info": (_i, [_p, _ip, C.POINTER(_i64)]),
    "femd_csr_download": (_i, [_p, _p, _p, _p]),
    "femd_csr_upload_values": (_i, [_p, _dp]),
    "femd_csr_device_arrays": (_i, [_p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p)]),
    "femd_csr_asymmetry": (_i, [_p, C.POINTER(_d)]),
    "femd_csr_diagonal": (_i, [_p, _dp]),
    "femd_csr_free": (None, [_p]),
    "femd_assemble_matrix": (_i, [_p, _i, _i, _ip, _ip, C.POINTER(Coefficient), _p]),
    "femd_assemble_vector": (_i, [_p, _i, _i, _ip, C.POINTER(Coefficient), _dp]),
    "femd_assemble_vector_dev": (_i, [_p, _i, _i, _ip, C.POINTER(Coefficient), _p, _p]),
    "femd_eval_coefficient": (_i, [_p, _i, C.POINTER(Coefficient), _dp]),
    "femd_assemble_matrix_add": (_i, [_p, _i, _i, _ip, _ip, C.POINTER(Coefficient), _p]),
    "femd_assemble_vector_add": (_i, [_p, _i, _i, _ip, C.POINTER(Coefficient), _dp]),
    "femd_facets_create": (_i, [_p, _i, _i, _ip, _i, C.POINTER(_p)]),
    "femd_facets_sizes": (_i, [_p, _ip]),
    "femd_facets_points": (_i, [_p, _dp, _dp, _dp, _dp, _dp]),
    "femd_facets_free": (None, [_p]),
    "femd_assemble_matrix_facets": (_i, [_p, _i, _ip, _ip, C.POINTER(Coefficient), _p, _i]),
    "femd_assemble_vector_facets": (_i, [_p, _i, _ip, C.POINTER(Coefficient), _dp, _i]),
    "femd_eval_coefficient_facets": (_i, [_p, C.POINTER(Coefficient), _dp]),
    "femd_spmv": (_i, [_p, _dp, _dp]),
    "femd_spmv_dev": (_i, [_p, _p, _p, _i, _p]),
    "femd_spmv_tuning": (_i, [_p, _i, C.c_char_p, _i, C.POINTER(_d)]),
    "femd_spmv_shape_count": (_i, []),
    "femd_spmv_shape_name": (_i, [_i, C.c_char_p, _i]),
    "femd_spmv_select": (_i, [_p, _i, C.c_char_p]),
    "femd_ddot": (_i, [_i64, _dp, _dp, C.POINTER(_d)]),
    "femd_ddot_dev": (_i, [_i64, _p, _p, C.POINTER(_d), _p]),
    "femd_precond_create": (_i, [_p, _i, _i, _d, C.POINTER(_p), C.POINTER(_d)]),
    "femd_precond_apply": (_i, [_p, _dp, _dp]),
    "femd_precond_apply_dev": (_i, [_p, _p, _p, _p]),
    "femd_precond_info": (_i, [_p, _ip]),
    "femd_precond_stats": (_i, [_p, _p]),
    "femd_rcm": (_i, [_i, _ip, _ip, _ip]),
    "femd_sweep_host_apply": (_i, [_i, _ip, _ip, _dp, _i, _d, _dp, _dp, _ip]),
    "femd_precond_free": (None, [_p]),
    "femd_cg": (_i, [_p, _p, _dp, _dp, _i, _d, C.POINTER(SolveInfo)]),
    "femd_cg_dev": (_i, [_p, _p, _p, _p, _i, _d, C.POINTER(SolveInfo), _p]),
    "femd_cg_dev_noalloc": (_i, [_p, _p, _p, _p, _i, _d, C.POINTER(SolveInfo), _p]),
    "femd_krylov": (_i, [_p, _p, _dp, _dp, _i, _i, _i, _d, _i, C.POINTER(SolveInfo)]),
    "femd_solve_refined": (_i, [_p, _dp, _dp, _d, _d, _i, C.POINTER(SolveInfo)]),
    "femd_solve_refined_dev": (_i, [_p, _p, _p, _d, _d, _i, C.POINTER(SolveInfo), _p]),
    "femd_batch_create": (_i, [_i, C.POINTER(_p), C.POINTER(_p)]),
    "femd_batch_info": (_i, [_p, C.POINTER(_i64), C.POINTER(_i64), _p]),
    "femd_batch_update_values": (_i, [_p]),
    "femd_batch_solve_refined": (_i, [_p, _dp, _dp, _d, _d, _i, _i, _p, _p, C.POINTER(SolveInfo)]),
    "femd_batch_solve_refined_dev": (_i, [_p, _p, _p, _d, _d, _i, _i, _p, _p, C.POINTER(SolveInfo), _p]),
    "femd_batch_solve": (_i, [_p, _dp, _dp, _d, _d, _i, _i, _i, _p, _p, C.POINTER(SolveInfo)]),
    "femd_batch_solve_dev": (_i, [_p, _p, _p, _d, _d, _i, _i, _i, _p, _p, C.POINTER(SolveInfo), _p]),
    "femd_batch_free": (None, [_p]),
    "femd_heat_run": (_i, [_p, _p, _dp, _i, _d, _i, _d, _i, _i, _i, _p, C.POINTER(_i), C.POINTER(SolveInfo)]),
    "femd_heat_run_dev": (_i, [_p, _p, _p, _i, _d, _i, _d, _i, _i, _i, _p, C.POINTER(_i), C.POINTER(SolveInfo), _p]),
    "femd_event_timer_start": (_i, [_p]),
    "femd_event_timer_stop": (_i, [_p, C.POINTER(_d)]),
    "femd_inner_timer": (_i, [_i]),
    "femd_inner_timer_read": (_i, [C.POINTER(_i64), C.POINTER(_d)]),
    "femd_inner_timer_shape": (_i, [C.c_char_p, _i]),
    "femd_inner_order": (_i, [_i]),
    "femd_assembly_mode": (_i, [_i]),
    "femd_assembly_kernel_time": (_i, [_p, _i, _i, _ip, _ip, C.POINTER(Coefficient), _i, _i, _dp, _dp]),
    "femd_inner_order_pattern": (_i, [_i, _ip, _ip, _i, C.c_void_p, _ip, _ip, _ip, _ip]),
    "femd_inner_order_pattern_gpu": (_i, [_i, _ip, _ip, _i, C.c_void_p, _ip, _ip, _ip, _ip]),
    "femd_device_setup": (_i, [_i]),
    "femd_pattern_lists": (_i, [C.c_void_p, _i] + [C.c_void_p] * 8),
    "femd_value_dictionary": (_i, [_i]),
    "femd_batch_dictionary": (_i, [_p, _i, C.POINTER(_i64), C.POINTER(_i)]),
    "femd_server_create": (_i, [_i64, _i, _i, _d, _d, _i, C.POINTER(_p)]),
    "femd_server_submit": (_i, [_p, _i64, _p, _p, _p, C.POINTER(_i64)]),
    "femd_server_step": (_i, [_p, _i, _p, _p, _p, _p, C.POINTER(_i)]),
    "femd_server_status": (_i, [_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i64)]),
    "femd_server_free": (None, [_p]),
}

_lib = None


def load():
    """The loaded library, its entry points typed.  Raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `python femd_amd/build.py` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


class FemdError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc == OK:
        return
    msg = load().femd_last_error().decode() or what
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == EZEROPIVOT:
        if "pivot" in msg:
            raise RuntimeError(msg)
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise FemdError(msg)


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def coefficient(prog=None, values=None, inputs=None):
    """A Coefficient struct and the arrays that keep its pointers alive.  inputs: the known fields
    the program reads, (space handle, coefficient vector, derivative code) per INPUT slot."""
    c = Coefficient()
    keep = []
    if values is not None:
        v = f64(values).reshape(-1)
        keep.append(v)
        c.values = v.ctypes.data_as(C.POINTER(_d))
        return c, keep
    code, nreg, consts = prog
    code = i32(code).reshape(-1)
    consts = f64(consts) if len(consts) else np.zeros(1)
    keep += [code, consts]
    c.ninstr = code.size // 4
    c.code = code.ctypes.data_as(C.POINTER(_i))
    c.nreg = int(nreg)
    c.nconst = int(len(prog[2]))
    c.consts = consts.ctypes.data_as(C.POINTER(_d))
    if inputs:
        fields = (Field * len(inputs))()
        for k, (space, vec, dcode) in enumerate(inputs):
            v = f64(vec).reshape(-1)
            keep.append(v)
            fields[k] = Field(space, v.ctypes.data_as(C.POINTER(_d)), int(dcode))
        keep.append(fields)
        c.ninputs = len(inputs)
        c.inputs = C.cast(fields, C.POINTER(Field))
    return c, keep


def coefficient_array(items):
    """items: list of (prog, values) or (prog, values, inputs); returns the ctypes array and the
    keep-alive list."""
    arr = (Coefficient * len(items))()
    keep = []
    for k, item in enumerate(items):
        c, kk = coefficient(*item)
        arr[k] = c
        keep += kk
    return arr, keep
