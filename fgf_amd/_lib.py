"""ctypes binding of libfgf_amd.so (include/fgf_amd.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation behind the Python mirror.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfgf_amd.so")

c_size = ctypes.c_size_t
c_u8p = ctypes.c_void_p       # device byte pointers travel as integers
c_vp = ctypes.c_void_p
c_int = ctypes.c_int
c_u64 = ctypes.c_uint64
c_plan = ctypes.c_void_p
c_ptrs = ctypes.POINTER(ctypes.c_void_p)
c_sizes = ctypes.POINTER(ctypes.c_size_t)

# name -> (restype, argtypes)
PROTOTYPES = {
    "fgf_last_error": (ctypes.c_char_p, []),
    "fgf_backend": (ctypes.c_char_p, []),
    "fgf_has_vector_elementwise": (c_int, [c_int]),
    "fgf_field_bytes": (c_size, [c_int]),
    "fgf_field_name": (ctypes.c_char_p, [c_int]),
    "fgf_elem_mul": (c_int, [c_int, c_u64, c_u64, ctypes.POINTER(c_u64)]),
    "fgf_elem_inv": (c_int, [c_int, c_u64, ctypes.POINTER(c_u64)]),
    "fgf_matrix_invert": (c_int, [c_int, c_size, c_vp, c_vp]),
    "fgf_plan_new": (c_int, [c_int, c_vp, c_size, ctypes.POINTER(c_plan)]),
    "fgf_plan_matrix": (c_int, [c_int, c_size, c_size, c_vp, c_size, ctypes.POINTER(c_plan)]),
    "fgf_coeff_new": (c_int, [c_int, c_u64, ctypes.POINTER(c_plan)]),
    "fgf_plan_free": (None, [c_plan]),
    "fgf_plan_len": (c_size, [c_plan]),
    "fgf_plan_dimensions": (None, [c_plan, c_sizes, c_sizes]),
    "fgf_plan_field": (c_int, [c_plan]),
    "fgf_plan_get": (c_int, [c_plan, c_size, ctypes.POINTER(c_u64)]),
    "fgf_plan_set_kernel": (c_int, [c_plan, c_int, c_int]),
    "fgf_plan_get_kernel": (c_int, [c_plan, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    "fgf_plan_get_at": (c_int, [c_plan, c_size, c_size, ctypes.POINTER(c_u64)]),
    "fgf_add_assign": (c_int, [c_int, c_u8p, c_size, c_u8p, c_size, c_vp]),
    "fgf_mul_add": (c_int, [c_int, c_u8p, c_size, c_u64, c_u8p, c_size, c_vp]),
    "fgf_mul_add_with": (c_int, [c_int, c_u8p, c_size, c_plan, c_size, c_u8p, c_size, c_vp]),
    "fgf_mul_into": (c_int, [c_int, c_u8p, c_size, c_u64, c_u8p, c_size, c_vp]),
    "fgf_mul_into_with": (c_int, [c_int, c_u8p, c_size, c_plan, c_size, c_u8p, c_size, c_vp]),
    "fgf_mul_assign": (c_int, [c_int, c_u8p, c_size, c_u64, c_vp]),
    "fgf_mul_assign_with": (c_int, [c_int, c_u8p, c_size, c_plan, c_size, c_vp]),
    "fgf_mul_add_scatter": (c_int, [c_int, c_u8p, c_size, c_size, c_vp, c_size, c_u8p, c_size, c_vp]),
    "fgf_mul_add_scatter_with": (c_int, [c_int, c_u8p, c_size, c_size, c_plan, c_u8p, c_size, c_vp]),
    "fgf_mul_add_gather": (c_int, [c_int, c_u8p, c_size, c_vp, c_size, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_mul_add_gather_with": (c_int, [c_int, c_u8p, c_size, c_plan, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_dot_product": (c_int, [c_int, c_u8p, c_size, c_vp, c_size, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_dot_product_with": (c_int, [c_int, c_u8p, c_size, c_plan, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_mul_add_matrix": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_ptrs, c_sizes, c_ptrs,
                                   c_sizes, c_size, c_vp]),
    "fgf_mul_add_matrix_with": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_plan, c_ptrs,
                                        c_sizes, c_size, c_vp]),
    "fgf_dot_product_matrix": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_ptrs, c_sizes,
                                       c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_dot_product_matrix_with": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_plan, c_ptrs,
                                            c_sizes, c_size, c_vp]),
    "fgf_stream_synchronize": (c_int, [c_vp]),
    "fgf_mul_add_matrix_scattered_with_host": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes,
                                                       c_size, c_plan, c_ptrs, c_sizes, c_size,
                                                       c_vp]),
    "fgf_dot_product_matrix_scattered_with_host": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes,
                                                           c_size, c_plan, c_ptrs, c_sizes, c_size,
                                                           c_vp]),
    "fgf_mul_add_matrix_with_host": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_plan,
                                             c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_dot_product_matrix_with_host": (c_int, [c_int, c_u8p, c_size, c_size, c_size, c_plan,
                                                 c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_mul_add_matrix_scattered": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes, c_size, c_ptrs,
                                             c_sizes, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_mul_add_matrix_scattered_noalloc": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes, c_size,
                                                     c_ptrs, c_sizes, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_mul_add_matrix_scattered_with": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes, c_size,
                                                  c_plan, c_ptrs, c_sizes, c_size, c_vp]),
    "fgf_dot_product_matrix_scattered_with": (c_int, [c_int, c_u8p, c_size, c_size, c_sizes,
                                                      c_size, c_plan, c_ptrs, c_sizes, c_size,
                                                      c_vp]),
    "fgf_mul_elementwise": (c_int, [c_int, c_u8p, c_size, c_u8p, c_size, c_u8p, c_size, c_vp]),
    "fgf_coeff_table": (c_int, [c_int, c_u64, ctypes.POINTER(ctypes.c_uint32)]),
    "fgf_batch_create": (c_int, [c_int, c_vp, c_size, ctypes.POINTER(c_vp)]),
    "fgf_batch_configure": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int]),
    "fgf_batch_run": (c_int, [c_vp, c_vp]),
    "fgf_batch_run_graph": (c_int, [c_vp, c_vp]),
    "fgf_batch_run_per_call": (c_int, [c_vp, c_vp]),
    "fgf_batch_items": (c_size, [c_vp]),
    "fgf_server_create": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "fgf_server_submit": (c_int, [c_vp, c_vp, ctypes.POINTER(c_u64)]),
    "fgf_server_drain": (c_int, [c_vp]),
    "fgf_server_result": (c_int, [c_vp, c_u64, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(c_int)]),
    "fgf_server_stats_get": (c_int, [c_vp, c_vp]),
    "fgf_server_cancel": (c_int, [c_vp, c_u64]),
    "fgf_server_wait": (c_int, [c_vp, c_u64, ctypes.c_double, ctypes.POINTER(c_int)]),
    "fgf_server_replay": (c_int, [c_vp, c_vp, c_vp, c_size, c_vp, ctypes.POINTER(ctypes.c_double)]),
    "fgf_server_destroy": (None, [c_vp]),
    "fgf_batch_bytes": (ctypes.c_uint64, [c_vp]),
    "fgf_batch_free": (None, [c_vp]),
    "fgf_pack": (c_int, [c_int, c_vp, c_size, ctypes.POINTER(c_u64), c_size]),
    "fgf_unpack": (c_int, [c_int, ctypes.POINTER(c_u64), c_size, c_vp, c_size]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m fgf_amd.build` (hipcc, gfx950)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib
