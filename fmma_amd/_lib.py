"""ctypes binding of libfmma_hip.so (the C-ABI declared in include/fmma_hip.h).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfmma_hip.so")

FMMA_OK = 0
FMMA_ERR_INVALID = 1
FMMA_ERR_SIZE = 2
FMMA_ERR_SOLVER = 3
FMMA_ERR_HIP = 4
FMMA_ERR_UNSUPPORTED = 5

FMMA_F64 = 0
FMMA_F32 = 1

KERNELS = {
    "inv_r": 0,
    "inv_r_reg": 1,
    "inv_diff": 2,
    "log_r": 3,
    "gauss": 4,
    "yukawa": 5,
    "sum_r2": 6,
    "sum2_r2": 7,
}

M2L_PRECISION = {"native": 0, "fp32": 1, "bf16x3": 2}

EXPORTS = [
    "fmma_plan_create", "fmma_plan_destroy", "fmma_plan_set_solver_type", "fmma_plan_set_poly_ord",
    "fmma_plan_set_depth", "fmma_plan_get_depth", "fmma_plan_set_trans_sym_flag", "fmma_plan_set_nrn_N",
    "fmma_plan_get_nrn_N",
    "fmma_plan_set_kernel", "fmma_plan_set_kernel_ops", "fmma_plan_set_m2l_precision", "fmma_plan_set_stream",
    "fmma_plan_set_timing", "fmma_plan_set_p2p_mutual", "fmma_plan_set_p2p_balance", "fmma_plan_gradient", "fmma_plan_gradient_host",
    "fmma_plan_solve_transpose", "fmma_plan_solve_transpose_host", "fmma_plan_solve_batch",
    "fmma_plan_set_batch_graph", "fmma_plan_set_m2l_autotune", "fmma_plan_set_m2l_tile", "fmma_plan_m2l_tiles", "fmma_plan_solve", "fmma_plan_solve_noalloc", "fmma_plan_solve_host", "fmma_plan_time_log",
    "fmma_plan_clear_time_log", "fmma_plan_level_coeffs", "fmma_plan_point_boxes", "fmma_morton_to_rowmajor",
    "fmma_last_error",
    "fmma_version",
]


class FMMAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fmma error %d: %s" % (code, msg))
        self.code = code


_lib = None
_hip = None


def pin_hip_runtime():
    """Load torch first when it is installed.  Its ROCm build brings its own
    libamdhip64 (and HSA runtime); with torch's loaded first, this library
    (and every Python kernel compiled by fmma_amd.jit) binds to that same
    runtime, so device pointers and streams pass between them.  Loaded in
    the other order, the process would hold two HIP runtimes and torch would
    find no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    pin_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ImportError("fmma_amd: %s is missing; build it with `python -m fmma_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, cp, i, l = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_long
    L.fmma_plan_create.argtypes = [ctypes.POINTER(vp), i, i]
    L.fmma_plan_destroy.argtypes = [vp]
    L.fmma_plan_set_solver_type.argtypes = [vp, cp]
    L.fmma_plan_set_poly_ord.argtypes = [vp, i]
    L.fmma_plan_set_depth.argtypes = [vp, i]
    L.fmma_plan_get_depth.argtypes = [vp, ctypes.POINTER(i)]
    L.fmma_plan_set_trans_sym_flag.argtypes = [vp, i]
    L.fmma_plan_set_nrn_N.argtypes = [vp, i]
    L.fmma_plan_get_nrn_N.argtypes = [vp, ctypes.POINTER(i)]
    L.fmma_plan_set_kernel.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double), i]
    L.fmma_plan_set_kernel_ops.argtypes = [vp, vp]
    L.fmma_plan_set_m2l_precision.argtypes = [vp, i]
    L.fmma_plan_set_stream.argtypes = [vp, vp]
    L.fmma_plan_set_timing.argtypes = [vp, i]
    L.fmma_plan_set_p2p_mutual.argtypes = [vp, i]
    L.fmma_plan_set_p2p_balance.argtypes = [vp, i]
    L.fmma_plan_gradient.argtypes = [vp, l, vp, vp, l, vp, vp, i]
    L.fmma_plan_gradient_host.argtypes = [vp, l, vp, vp, l, vp, vp]
    L.fmma_plan_solve_transpose.argtypes = [vp, l, vp, vp, l, vp, vp]
    L.fmma_plan_solve_transpose_host.argtypes = [vp, l, vp, vp, l, vp, vp]
    lp = ctypes.POINTER(ctypes.c_long)
    L.fmma_plan_solve_batch.argtypes = [vp, i, lp, vp, vp, lp, vp, vp]
    L.fmma_plan_set_batch_graph.argtypes = [vp, i]
    L.fmma_plan_set_m2l_autotune.argtypes = [vp, i]
    L.fmma_plan_set_m2l_tile.argtypes = [vp, i]
    L.fmma_plan_m2l_tiles.argtypes = [vp, i, lp, ctypes.POINTER(ctypes.c_int), i]
    L.fmma_plan_solve.argtypes = [vp, l, vp, vp, l, vp, vp]
    L.fmma_plan_solve_noalloc.argtypes = [vp, l, vp, vp, l, vp, vp]
    L.fmma_plan_solve_host.argtypes = [vp, l, vp, vp, l, vp, vp]
    L.fmma_plan_time_log.argtypes = [vp, ctypes.POINTER(cp), ctypes.POINTER(ctypes.c_double), i]
    L.fmma_plan_clear_time_log.argtypes = [vp]
    L.fmma_plan_level_coeffs.argtypes = [vp, i, i, ctypes.POINTER(vp), ctypes.POINTER(l), ctypes.POINTER(i)]
    L.fmma_plan_point_boxes.argtypes = [vp, i, l, ctypes.POINTER(l)]
    L.fmma_morton_to_rowmajor.argtypes = [i, i, l]
    L.fmma_morton_to_rowmajor.restype = l
    L.fmma_last_error.restype = cp
    L.fmma_version.restype = cp
    _lib = L
    return L


def check(rc):
    if rc != FMMA_OK:
        raise FMMAError(rc, lib().fmma_last_error().decode())
    return rc


def hip():
    """libamdhip64 for the few raw copies the wrapper needs (device -> host)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def memcpy_d2h(dst_ptr, src_ptr, nbytes):
    rc = hip().hipMemcpy(dst_ptr, src_ptr, nbytes, 2)
    if rc != 0:
        raise FMMAError(FMMA_ERR_HIP, "hipMemcpy D2H failed (%d)" % rc)
