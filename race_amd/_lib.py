"""ctypes binding of librace_amd.so (include/race_amd.h).

The library is built in-tree by race_amd.build; loading fails loudly when it
is missing -- there is no CPU substitute for the GPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librace_amd.so")

SUCCESS = 0
STATUS_NAMES = {
    0: "RACE_SUCCESS", 1: "RACE_ERR_INVALID_ARG", 2: "RACE_ERR_MATRIX_SYMM", 3: "RACE_ERR_D2_COLOR",
    4: "RACE_ERR_D1_COLOR", 5: "RACE_ERR_GRAPH_TRAVERSAL", 6: "RACE_ERR_INCOMPATIBILITY",
    7: "RACE_ERR_HWLOC", 8: "RACE_ERR_PIN", 9: "RACE_ERR_INTERNAL", 10: "RACE_ERR_NOT_IMPLEMENTED",
}

ONE, TWO, POWER = 1, 2, 3
TWO_BLOCK, THREE_BLOCK = 0, 1
LB_ROW, LB_NNZ = 0, 1
FILL, SCATTER = 0, 1
K_SYMMSPMV, K_GS, K_KACZ, K_SPMV, K_MPK, K_SPMTV, K_MPK_HIGHEST = 0, 1, 2, 3, 4, 5, 6
FP64, FP32, BF16 = 0, 1, 2

i32p = ctypes.POINTER(ctypes.c_int)
f64p = ctypes.POINTER(ctypes.c_double)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p


class RaceError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")


class KernelDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("nrow", ctypes.c_int), ("rowPtr", i32p), ("col", i32p),
        ("val", f64p), ("val_precision", ctypes.c_int), ("vec_precision", ctypes.c_int),
        ("nrhs", ctypes.c_int), ("x", vp), ("b", vp), ("power", ctypes.c_int),
        ("sub_power", ctypes.c_int),
    ]


COLOR_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
POWER_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_void_p)
COMM_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
COMM_SIMPLE_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p)

# name -> (restype, argtypes)
PROTOTYPES = {
    "race_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p,
                                   i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p,
                                   ctypes.c_int, ctypes.c_int]),
    "race_destroy": (ctypes.c_int, [vp]),
    "race_color": (ctypes.c_int, [vp]),
    "race_color_power": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_char, ctypes.c_int]),
    "race_get_perm": (ctypes.c_int, [vp, i32p, ctypes.c_int]),
    "race_get_inv_perm": (ctypes.c_int, [vp, i32p, ctypes.c_int]),
    "race_get_num_threads": (ctypes.c_int, [vp, i32p]),
    "race_get_efficiency": (ctypes.c_int, [vp, f64p]),
    "race_get_max_stage_depth": (ctypes.c_int, [vp, i32p]),
    "race_get_highest_power": (ctypes.c_int, [vp, i32p]),
    "race_register_function": (ctypes.c_int, [vp, COLOR_FN, vp, i32p]),
    "race_attach_communication": (ctypes.c_int, [vp, ctypes.c_int, COMM_FN, vp]),
    "race_attach_communication_simple": (ctypes.c_int, [vp, ctypes.c_int, COMM_SIMPLE_FN, vp]),
    "race_register_power_function": (ctypes.c_int, [vp, POWER_FN, vp, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, i32p]),
    "race_register_device": (ctypes.c_int, [vp, ctypes.POINTER(KernelDesc), i32p]),
    "race_set_vectors": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
    "race_set_option": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
    "race_register_refinement": (ctypes.c_int, [vp, ctypes.POINTER(KernelDesc), ctypes.c_int, ctypes.c_int,
                                                i32p]),
    "race_refine": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, i32p,
                                   ctypes.POINTER(ctypes.c_long), f64p]),
    "race_residual": (ctypes.c_int, [vp, ctypes.c_int, f64p]),
    "race_refine_noalloc": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long),
                                           ctypes.POINTER(ctypes.c_double)]),
    "race_refine_noalloc_max_bytes": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "race_execute": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
    "race_execute_async": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp]),
    "race_set_power": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
    "race_get_power": (ctypes.c_int, [vp, ctypes.c_int, i32p]),
    "race_tune_function": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i32p]),
    "race_check_power_coverage": (ctypes.c_int, [vp]),
    "race_set_sub_power": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
    "race_get_sub_power": (ctypes.c_int, [vp, ctypes.c_int, i32p]),
    "race_set_serial": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
    "race_reset_time": (ctypes.c_int, [vp]),
    "race_sleep": (ctypes.c_int, [vp]),
    "race_pin_thread": (ctypes.c_int, [vp, ctypes.c_int]),
    "race_numa_init_row_ptr": (ctypes.c_int, [vp, i32p]),
    "race_numa_init_row_ptr_local": (ctypes.c_int, [vp, ctypes.POINTER(i32p)]),
    "race_numa_init_mtx_vec": (ctypes.c_int, [vp, i32p, i32p, f64p, f64p, ctypes.c_int]),
    "race_numa_init_mtx_vec_local": (ctypes.c_int, [vp, ctypes.POINTER(i32p), ctypes.POINTER(i32p),
                                                    ctypes.POINTER(f64p), ctypes.c_int]),
    "race_get_numa_splitting": (ctypes.c_int, [vp, i32p, ctypes.c_int, i32p]),
    "race_get_schedule": (ctypes.c_int, [vp, ctypes.c_int, i32p, i32p, i32p, i32p, i32p, ctypes.c_int]),
    "race_get_tree": (ctypes.c_int, [vp, i32p, i32p, ctypes.c_int]),
    "race_get_mpk_groups": (ctypes.c_int, [vp, i32p, i32p, ctypes.c_int]),
    "race_get_launch_count": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i32p]),
    "race_get_flow_aborts": (ctypes.c_int, [vp, ctypes.c_int, i32p]),
    "race_rcm": (ctypes.c_int, [ctypes.c_int, i32p, i32p, i32p, i32p]),
    "race_schedule_save": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "race_schedule_load": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "race_schedule_info": (ctypes.c_int, [vp, ctypes.c_size_t, i32p, i32p, i32p]),
    "race_get_wave_count": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i32p]),
    "race_get_preprocess_stats": (ctypes.c_int, [vp, f64p, i32p]),
    "race_timer_begin": (ctypes.c_int, [vp]),
    "race_timer_end": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float)]),
    "race_lb_split": (ctypes.c_int, [i32p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p,
                                     i32p, i32p, i32p]),
    "race_split_threads": (ctypes.c_int, [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]),
    "race_mpk_hopeless_regions": (ctypes.c_int, [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p,
                                                 ctypes.c_int, i32p]),
    "race_mpk_consolidate": (ctypes.c_int, [i32p, f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double, i32p,
                                            ctypes.c_int, i32p, i32p, i32p, ctypes.c_int, i32p, i32p]),
    "race_get_mpk_children": (ctypes.c_int, [vp, i32p, i32p, ctypes.c_int]),
    "race_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]),
    "race_batch_add": (ctypes.c_int, [vp, vp, ctypes.POINTER(KernelDesc)]),
    "race_batch_set_option": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
    "race_batch_get_option": (ctypes.c_int, [vp, ctypes.c_char_p, i32p]),
    "race_batch_get_launch_count": (ctypes.c_int, [vp, ctypes.c_int, i32p]),
    "race_batch_get_flow_aborts": (ctypes.c_int, [vp, i32p]),
    "race_batch_build": (ctypes.c_int, [vp]),
    "race_batch_execute": (ctypes.c_int, [vp, ctypes.c_int]),
    "race_batch_execute_async": (ctypes.c_int, [vp, ctypes.c_int, vp]),
    "race_batch_time": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "race_batch_candidates": (ctypes.c_int, [ctypes.c_int, i32p]),
    "race_batch_candidate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p]),
    "race_batch_use_candidate": (ctypes.c_int, [vp, ctypes.c_int]),
    "race_batch_autotune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i32p,
                                           ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "race_batch_destroy": (ctypes.c_int, [vp]),
    "race_device_copy_async": (ctypes.c_int, [vp, vp, ctypes.c_size_t, vp]),
    "race_get_stream": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "race_mlp_dims": (ctypes.c_int, [i32p, i32p, i32p, i32p]),
    "race_matrix_features": (ctypes.c_int, [ctypes.c_int, i32p, i32p, f64p]),
    "race_mlp_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_uint]),
    "race_mlp_set_params": (ctypes.c_int, [vp, f32p, ctypes.c_int]),
    "race_mlp_get_params": (ctypes.c_int, [vp, f32p, ctypes.c_int]),
    "race_mlp_forward": (ctypes.c_int, [vp, f32p, ctypes.c_int, f32p, ctypes.c_int]),
    "race_mlp_gradients": (ctypes.c_int, [vp, f32p, f32p, f32p, ctypes.c_int, f32p,
                                          ctypes.POINTER(ctypes.c_float)]),
    "race_mlp_train": (ctypes.c_int, [vp, f32p, f32p, f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.POINTER(ctypes.c_float)]),
    "race_mlp_destroy": (ctypes.c_int, [vp]),
    "race_sell_c_sigmize": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p]),
    "race_simdify": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p, vp,
                                    ctypes.c_int, ctypes.c_int]),
    "race_simdify_resolve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, vp,
                                            ctypes.c_int, ctypes.c_int]),
    "race_simdify_d1": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, vp,
                                       ctypes.c_int, ctypes.c_int]),
    "race_device_malloc": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t]),
    "race_device_free": (ctypes.c_int, [vp]),
    "race_memcpy_h2d": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "race_memcpy_d2h": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "race_memset_device": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_size_t]),
    "race_device_synchronize": (ctypes.c_int, []),
    "race_device_copy": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    "race_device_count": (ctypes.c_int, [i32p]),
    "race_error_string": (ctypes.c_char_p, [ctypes.c_int]),
}

_lib = None


def load():
    """Load librace_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m race_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status, what):
    if status != SUCCESS:
        raise RaceError(status, what)
