"""ctypes binding of the C ABI declared in include/mlip_hip.h."""
import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libmlip_hip.so")

MLIP_OK = 0
ERRORS = {1: "argument", 2: "file", 3: "model", 4: "device", 5: "state", 6: "full"}
MLIP_ERR_FULL = 6


class MlipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mlip_hip %s error: %s" % (ERRORS.get(code, str(code)), msg))
        self.code = code


_lib = None


def build_library():
    """Compile the HIP engine for gfx950 (in-tree, into _build/)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR], check=True)


def library():
    """Load the HIP engine.  There is no fallback: without the library the
    path is unavailable and this raises."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MlipError(4, "HIP library %s is not built (run build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip, i64p = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_int64))
    c_int, c_double, c_char_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t
    L.mlip_create.argtypes = [c_char_p, c_int, ctypes.POINTER(c_char_p), ctypes.POINTER(vp), c_char_p, c_size_t]
    L.mlip_create.restype = c_int
    L.mlip_destroy.argtypes = [vp]
    L.mlip_destroy.restype = None
    L.mlip_cutoff.argtypes = [vp]
    L.mlip_cutoff.restype = c_double
    L.mlip_masses.argtypes = [vp, dp]
    L.mlip_settings.argtypes = [c_int]
    L.mlip_set_precision.argtypes = [vp, c_int]
    L.mlip_set_atom_mode.argtypes = [vp, c_int]
    L.mlip_last_error.argtypes = [vp]
    L.mlip_last_error.restype = c_char_p
    L.mlip_compute.argtypes = [vp, c_int, c_int, dp, ip, i64p, c_int, ip, ip, ctypes.POINTER(ip), c_int, c_int,
                               c_int, dp, dp, dp, dp, dp]
    L.mlip_compute_flat.argtypes = [vp, c_int, c_int, dp, ip, i64p, c_int, ip, ip, ip, ip, c_int, c_int, c_int,
                                    dp, dp, dp, dp, dp]
    L.mlip_compute_flat_noalloc.argtypes = L.mlip_compute_flat.argtypes
    L.mlip_resident_set.argtypes = [vp, c_int, dp, ip, ctypes.c_int64, i64p, ip, dp]
    L.mlip_resident_update_positions.argtypes = [vp, dp]
    L.mlip_resident_compute.argtypes = [vp]
    L.mlip_resident_sync.argtypes = [vp]
    L.mlip_resident_get.argtypes = [vp, dp, dp, dp, dp]
    L.mlip_resident_set_cell.argtypes = [vp, c_int, dp, dp, ip]
    L.mlip_resident_npairs.argtypes = [vp]
    L.mlip_resident_npairs.restype = ctypes.c_int64
    L.mlip_resident_lists.argtypes = [vp, i64p, ip, dp, ip, ip]
    L.mlip_set_neighbor_builder.argtypes = [vp, c_int]
    L.mlip_train_features.argtypes = [vp, dp, dp, dp, dp]
    L.mlip_profile_enable.argtypes = [vp, c_int]
    L.mlip_profile_read.argtypes = [vp, ctypes.POINTER(c_char_p), dp, ip, c_int]
    L.mlip_flops.argtypes = [vp, dp]
    L.mlip_random_structures.argtypes = [c_int, c_int, c_int, c_double, c_double, ctypes.c_uint64, ip, dp, dp, ip,
                                         ctypes.c_int64]
    L.mlip_set_q34_plan.argtypes = [vp, c_int, c_int, c_int]
    L.mlip_set_force_threads.argtypes = [vp, c_int]
    L.mlip_flops_reference.argtypes = [vp, dp]
    i64 = ctypes.c_int64
    L.mlip_pool_bytes.argtypes = [vp, i64, i64, c_int]
    L.mlip_pool_bytes.restype = i64
    L.mlip_pool_init.argtypes = [vp, i64, i64, c_int]
    L.mlip_pool_admit.argtypes = [vp, c_int, dp, dp, ip, dp, c_int, c_double, c_double, ip]
    L.mlip_pool_admit_many.argtypes = [vp, c_int, ip, dp, dp, ip, dp, ip, c_double, dp, c_int, ip]
    L.mlip_pool_tick.argtypes = [vp, ip]
    L.mlip_pool_sync.argtypes = [vp, ip]
    L.mlip_pool_done.argtypes = [vp, ip, c_int]
    L.mlip_pool_result.argtypes = [vp, c_int, dp, dp, dp, dp, dp, dp, c_int]
    L.mlip_pool_stats.argtypes = [vp, i64p]
    L.mlip_pool_rebuilds.argtypes = [vp]
    L.mlip_pool_rebuilds.restype = i64
    L.mlip_batch_set.argtypes = [vp, c_int, ip, dp, dp, ip]
    L.mlip_batch_get.argtypes = [vp, dp, dp, dp]
    L.mlip_graph_mode.argtypes = [vp, c_int]
    L.mlip_autotune.argtypes = [vp, c_int, ip]
    L.mlip_autotune_q34.argtypes = [vp, c_int, ip]
    L.mlip_set_anlm_kernel.argtypes = [vp, c_int]
    L.mlip_get_anlm_kernel.argtypes = [vp]
    L.mlip_calibration_copy.argtypes = [vp, ctypes.c_int64, c_int]
    L.mlip_debug_fetch.argtypes = [vp, c_int, dp, ctypes.c_int64]
    L.mlip_debug_fetch.restype = ctypes.c_int64
    L.mlip_model_info.argtypes = [vp, ip]
    L.mlip_nlist_build.argtypes = [dp, c_int, dp, c_double, ctypes.POINTER(vp), c_char_p, c_size_t]
    L.mlip_nlist_npairs.argtypes = [vp]
    L.mlip_nlist_npairs.restype = ctypes.c_int64
    L.mlip_nlist_copy.argtypes = [vp, i64p, ip, dp]
    L.mlip_nlist_free.argtypes = [vp]
    L.mlip_nlist_free.restype = None
    L.mlip_model_inspect.argtypes = [c_char_p, c_int, ip, ctypes.c_int64, c_char_p, c_size_t]
    _lib = L
    return L
