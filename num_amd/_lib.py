"""ctypes binding of libnum_amd.so (the C ABI declared in include/num_amd.h).

The library is the only compute path: if it is missing, or a call fails, an
exception is raised -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnum_amd.so")

NUM_OK, NUM_EINVAL, NUM_ERUNTIME, NUM_EHIP, NUM_ENOMEM = 0, 1, 2, 3, 4

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f64 = C.c_double
pi32 = C.POINTER(C.c_int32)
pi64 = C.POINTER(C.c_int64)


class CgResult(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("breakdown", C.c_int32),
                ("pad_", C.c_int32), ("residual", C.c_double)]


class MixedResult(C.Structure):
    _fields_ = [("outer", C.c_int32), ("inner_total", C.c_int32), ("converged", C.c_int32),
                ("breakdown", C.c_int32), ("true_residual", C.c_double)]


# name -> (restype, argtypes); every symbol include/num_amd.h declares.
SIGNATURES = {
    "num_init": (C.c_int, [i32]),
    "num_finalize": (C.c_int, []),
    "num_last_error": (C.c_char_p, []),
    "num_abi_version": (i32, []),
    "num_csr_create": (C.c_int, [vp, vp, vp, vp, i32, i32, C.POINTER(vp)]),
    "num_csr_create_host": (C.c_int, [vp, vp, vp, i32, i32, C.POINTER(vp)]),
    "num_csr_attach_f32": (C.c_int, [vp, vp]),
    "num_csr_attach_rel16": (C.c_int, [vp, vp, vp, pi32]),
    "num_csr_attach_d16": (C.c_int, [vp, vp, vp, pi32]),
    "num_csr_index_info": (C.c_int, [vp, pi32]),
    "num_csr_set_windows": (C.c_int, [vp, i32, vp]),
    "num_csr_window_info": (C.c_int, [vp, pi32]),
    "num_csr_set_mixed_path": (C.c_int, [vp, i32]),
    "num_csr_pattern_info": (C.c_int, [vp, pi32, pi32]),
    "num_csr_attach_xcd": (C.c_int, [vp, vp, i32, vp, vp, vp]),
    "num_spmv_xcd": (C.c_int, [vp, vp, vp, vp, vp]),
    "num_csr_destroy": (C.c_int, [vp]),
    "num_csr_shape": (C.c_int, [vp, pi32, pi32, pi64]),
    "num_csr_set_schedule": (C.c_int, [vp, i32, i32, i32, i32, i32]),
    "num_csr_schedule_info": (C.c_int, [vp, pi32, pi32, pi32, pi32, pi32, pi32, pi32, pi32]),
    "num_spmv": (C.c_int, [vp, vp, vp, vp]),
    "num_spmv_rows": (C.c_int, [vp, i32, i32, vp, vp, vp]),
    "num_spmv_split": (C.c_int, [vp, i32, i32, vp, vp, vp, vp]),
    "num_spmv_f32": (C.c_int, [vp, vp, vp, vp]),
    "num_spmv_host": (C.c_int, [vp, vp, vp, i32, i32, vp, vp]),
    "num_batch_spmv": (C.c_int, [vp, vp, vp, i32, vp]),
    "num_batch_spmv_host": (C.c_int, [vp, vp, vp, i32, i32, i32, vp, vp]),
    "num_axpy": (C.c_int, [i64, f64, vp, vp, vp, vp]),
    "num_axpy2": (C.c_int, [i64, f64, vp, f64, vp, vp, vp]),
    "num_scale": (C.c_int, [i64, f64, vp, vp]),
    "num_fill": (C.c_int, [i64, f64, vp, vp]),
    "num_copy": (C.c_int, [i64, vp, vp, vp]),
    "num_dot_workspace_bytes": (i64, [i64]),
    "num_dot": (C.c_int, [i64, vp, vp, vp, vp, vp]),
    "num_dot_noalloc": (C.c_int, [i64, vp, vp, vp, vp]),
    "num_norm2": (C.c_int, [i64, vp, vp, vp, vp]),
    "num_cg_workspace_bytes": (i64, [vp]),
    "num_cg_solve": (C.c_int, [vp, vp, vp, f64, i32, vp, i64, C.POINTER(CgResult), vp]),
    "num_cg_mixed_workspace_bytes": (i64, [vp]),
    "num_cg_solve_mixed": (C.c_int, [vp, vp, vp, f64, i32, f64, vp, i64,
                                     C.POINTER(MixedResult), vp]),
    "num_engine_create": (C.c_int, [i32, i32, i32, C.POINTER(vp)]),
    "num_engine_destroy": (C.c_int, [vp]),
    "num_engine_work_bytes": (i64, [i32, i32]),
    "num_engine_add": (C.c_int, [vp, vp, vp, vp, f64, i32, i32, f64, vp, i64, vp, pi32]),
    "num_engine_stencil_bytes": (i64, [i32, i32, i32, i32, i32]),
    "num_engine_add_stencil": (C.c_int, [vp, i32, i32, i32, i32, f64, C.c_uint64, f64, i32, i32,
                                         f64, vp, i64, vp, pi32]),
    "num_engine_set_graph": (C.c_int, [vp, i32]),
    "num_engine_set_schedule": (C.c_int, [vp, i32]),
    "num_engine_launch": (C.c_int, [vp, i32, vp]),
    "num_engine_poll": (C.c_int, [vp, i32, pi32, pi32]),
    "num_engine_step": (C.c_int, [vp, i32, pi32, vp]),
    "num_engine_result": (C.c_int, [vp, i32, C.POINTER(CgResult), pi32, vp]),
    "num_engine_finished": (C.c_int, [vp, pi32, i32, pi32]),
    "num_engine_remove": (C.c_int, [vp, i32]),
    "num_engine_info": (C.c_int, [vp, pi32, pi32, pi32]),
    "num_stencil_fill": (C.c_int, [i32, i32, i32, i32, f64, vp, vp, vp, vp]),
    "num_hash_uniform": (C.c_int, [i64, C.c_uint64, vp, vp]),
    "num_halo_pack": (C.c_int, [vp, i32, vp, vp, vp]),
    "num_halo_scatter": (C.c_int, [vp, i32, vp, i32, vp]),
}

_lib = None


def load():
    """Load libnum_amd.so (built by ``python -m num_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m num_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    m = load().num_last_error()
    return m.decode(errors="replace") if m else ""


def check(rc: int, what: str = "") -> None:
    """Map a NUM_E* code to the Python exception the reference raises."""
    if rc == NUM_OK:
        return
    msg = last_error() or f"error code {rc}"
    if what:
        msg = f"{what}: {msg}"
    if rc == NUM_EINVAL:
        raise ValueError(msg)
    if rc == NUM_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
