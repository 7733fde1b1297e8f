"""ctypes loader of libefg.so (the HIP library).  There is no CPU fallback:
importing this module without the built library raises."""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libefg.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "efg.h")

c_f32p = ctypes.POINTER(ctypes.c_float)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_sz = ctypes.c_size_t
c_szp = ctypes.POINTER(ctypes.c_size_t)
c_int = ctypes.c_int
c_intp = ctypes.POINTER(ctypes.c_int)
c_vp = ctypes.c_void_p


class BuildParams(ctypes.Structure):
    _fields_ = [("trees", c_int), ("trees_build", c_int), ("merge_level", c_int), ("epochs", c_int),
                ("check", c_int), ("L", c_int), ("K", c_int), ("S", c_int), ("tns", c_int),
                ("seed", ctypes.c_uint64)]


_SIGS = {
    "efg_default_build_params": (None, [ctypes.POINTER(BuildParams), c_int]),
    "efg_last_error": (ctypes.c_char_p, []),
    "efg_create": (c_int, [c_f32p, c_sz, c_sz, c_sz, ctypes.POINTER(BuildParams), ctypes.POINTER(c_vp)]),
    "efg_destroy": (c_int, [c_vp]),
    "efg_build_index": (c_int, [c_vp]),
    "efg_build_trees": (c_int, [c_vp]),
    "efg_set_search_params": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int]),
    "efg_get_search_params": (c_int, [c_vp, c_intp, c_intp, c_intp, c_intp, c_intp]),
    "efg_set_search_precision": (c_int, [c_vp, c_int]),
    "efg_set_build_precision": (c_int, [c_vp, c_int]),
    "efg_set_join_staging_margin": (c_int, [c_vp, ctypes.c_double]),
    "efg_set_join_routing": (c_int, [c_vp, c_int]),
    "efg_hash_set_codes": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_u64p, c_u64p, c_sz]),
    "efg_hash_load_codes": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, ctypes.c_char_p, ctypes.c_char_p]),
    "efg_knn_search": (c_int, [c_vp, c_int, c_f32p, c_sz, c_sz]),
    "efg_get_results": (c_int, [c_vp, c_szp, c_intp, c_i32p, c_f32p]),
    "efg_knn_search_device": (c_int, [c_vp, c_int, c_vp, c_sz, c_sz, c_vp, c_vp, c_vp, c_vp]),
    "efg_knn_search_device_noalloc": (c_int, [c_vp, c_int, c_vp, c_sz, c_sz, c_vp, c_vp, c_vp]),
    "efg_graph_size": (c_int, [c_vp, c_szp]),
    "efg_graph_row": (c_int, [c_vp, c_sz, c_u32p, c_sz, c_szp]),
    "efg_save_graph": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_load_graph": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_save_trees": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_load_trees": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_save_index": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_load_index": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_save_results": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_save_image": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_load_image": (c_int, [c_vp, ctypes.c_char_p]),
    "efg_exact_knn_device": (c_int, [c_vp, c_sz, c_sz, c_sz, c_vp, c_sz, c_sz, c_int, c_int, c_vp, c_vp, c_vp]),
    "efg_exact_knn_device_prec": (c_int, [c_vp, c_sz, c_sz, c_sz, c_vp, c_sz, c_sz, c_int, c_int, c_int, c_vp, c_vp,
                                          c_vp]),
    "efg_tree_shape": (c_int, [c_vp, c_intp, c_szp]),
    "efg_get_trees": (c_int, [c_vp, c_i32p, c_f32p, c_i32p, c_i32p, c_u32p, c_u32p, c_i32p, c_u32p]),
    "efg_set_trees": (c_int, [c_vp, c_int, c_sz, c_i32p, c_f32p, c_i32p, c_i32p, c_u32p, c_u32p, c_i32p, c_u32p]),
    "efg_get_graph": (c_int, [c_vp, c_u64p, c_u32p, c_sz]),
    "efg_set_graph": (c_int, [c_vp, c_u64p, c_u32p]),
    "efg_device_data": (c_int, [c_vp, ctypes.POINTER(c_vp), c_szp]),
    "efg_get_counters": (c_int, [c_vp, c_u64p, c_int]),
    "efg_get_join_pairs": (c_int, [c_vp, c_u64p, c_int, c_intp]),
    "efg_build_stage": (c_int, [c_vp, c_int]),
    "efg_get_pools": (c_int, [c_vp, c_u64p, c_u8p, c_u32p, c_u32p]),
    "efg_get_lists": (c_int, [c_vp, c_int, c_u64p, c_u32p, c_sz]),
    "efg_get_build_times": (c_int, [c_vp, ctypes.POINTER(ctypes.c_double), c_int]),
    "efg_set_profiling": (c_int, [c_vp, c_int]),
    "efg_get_kernel_times": (c_int, [c_vp, ctypes.POINTER(ctypes.c_double), c_int, c_intp]),
    "efg_get_search_work": (c_int, [c_vp, c_u64p, c_int]),
    "efg_knn_search_batch": (c_int, [c_vp, c_f32p, c_sz, c_sz, c_intp, c_intp, c_intp, c_int, c_i32p, c_f32p, c_int]),
    "efg_knn_search_batch_device": (c_int, [c_vp, c_vp, c_sz, c_sz, c_intp, c_intp, c_intp, c_int, c_vp, c_vp, c_int,
                                            c_vp]),
    "efg_batch_plan_create": (c_int, [c_vp, c_vp, c_sz, c_sz, c_intp, c_intp, c_intp, c_int, c_vp, c_vp, c_int,
                                      c_vp, ctypes.POINTER(c_vp)]),
    "efg_batch_plan_run": (c_int, [c_vp, c_vp]),
    "efg_batch_plan_destroy": (c_int, [c_vp]),
    "efg_set_search_threads": (c_int, [c_vp, c_int]),
    "efg_autotune_search": (c_int, [c_vp, c_int, c_f32p, c_sz, c_sz, c_intp]),
    "efg_mlp_param_count": (c_int, []),
    "efg_mlp_init": (c_int, [c_f32p, ctypes.c_uint64]),
    "efg_mlp_logits": (c_int, [c_f32p, c_f32p, c_f32p, c_sz, c_f32p]),
    "efg_mlp_gradient": (c_int, [c_f32p, c_f32p, c_f32p, c_u8p, c_sz, c_f32p, c_f32p]),
    "efg_mlp_train": (c_int, [c_f32p, c_u8p, c_sz, c_f32p, c_f32p, c_int, c_int, ctypes.c_float, ctypes.c_uint64,
                              c_f32p]),
    "efg_set_probe_params": (c_int, [c_vp, c_int, c_int, c_int]),
    "efg_search_traces": (c_int, [c_vp, c_int, c_f32p, c_sz, c_sz, c_i32p, c_f32p, c_u8p]),
    "efg_set_stopping_network": (c_int, [c_vp, c_f32p, c_f32p, ctypes.c_float]),
    "efg_set_search_mode": (c_int, [c_vp, c_int]),
    "efg_get_learned_stats": (c_int, [c_vp, c_u64p, c_u64p, c_u64p]),
    "efg_get_learned_restarts": (c_int, [c_vp, c_u64p]),
    "efg_set_batch_form": (c_int, [c_vp, c_int]),
    "efg_get_batch_shape": (c_int, [c_vp, c_intp, c_intp, c_intp, c_intp]),
}

_lib = None


class EfgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"efg error {code}: {msg}")
        self.code = code
        self.msg = msg


def header_symbols() -> list[str]:
    """Function names declared in include/efg.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(efg_\w+)\s*\(", txt, re.M)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libefg.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise EfgError(rc, lib().efg_last_error().decode())
