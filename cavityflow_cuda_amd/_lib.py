"""ctypes binding of libcavityflow_hip.so, the C ABI declared in include/lbm_abi.h.

The product path has no CPU fallback: importing this module fails loudly when
the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcavityflow_hip.so")

# enum lbm_flag / lbm_direction / lbm_precision / lbm_status
OBSTACLE, FLUID, VELOCITY_INJECTION, GHOST_LAYER = 1, 2, 4, 8
LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT = range(6)
F32, F64, F16S, BF16S, F16C = 0, 1, 2, 3, 4
LAUNCH_NONTEMPORAL, LAUNCH_XCD_REMAP, LAUNCH_WAVES4, LAUNCH_WAVES3, LAUNCH_WAVES2 = 1, 2, 4, 8, 16
LAUNCH_ADDR64 = 32
LAUNCH_WIDE = 64
LAUNCH_LDS = 128
LAUNCH_BIG = 256
LBM_OK = 0
NOALLOC_MAX_BOXES = 16  # LBM_NOALLOC_MAX_BOXES
(LBM_ERR_ARG, LBM_ERR_BOX, LBM_ERR_HIP, LBM_ERR_NOMEM, LBM_ERR_PARAMS, LBM_ERR_STATE,
 LBM_ERR_UNSUPPORTED) = range(-1, -8, -1)

EXPORTED = (
    "lbm_compute_parameters", "lbm_create", "lbm_destroy", "lbm_get_device_arrays",
    "lbm_simulation_step_alpha", "lbm_simulation_step_beta",
    "lbm_get_density_distributions", "lbm_set_density_distributions",
    "lbm_set_density_distributions_dir", "lbm_get_flags", "lbm_set_flags",
    "lbm_get_velocities", "lbm_set_velocities", "lbm_get_densities", "lbm_set_densities",
    "lbm_face_box", "lbm_face_elements", "lbm_pack_face", "lbm_unpack_face",
    "lbm_unpack_face_pair", "lbm_pack_face_pair",
    "lbm_exchange_face", "lbm_exchange_faces", "lbm_device_alloc", "lbm_device_free",
    "lbm_set_launch_options", "lbm_get_launch_options", "lbm_set_step_pack",
    "lbm_set_step_unpack", "lbm_get_moments",
    "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_size", "lbm_batch_set_tuning",
    "lbm_batch_get_tuning", "lbm_batch_step", "lbm_batch_run", "lbm_batch_launch",
    "lbm_batch_autotune",
    "lbm_batch_residual", "lbm_reset", "lbm_device_bytes", "lbm_state_bytes",
    "lbm_save_state", "lbm_load_state",
    "lbm_reduce_box", "lbm_reduce_boxes_noalloc", "lbm_residual", "lbm_sync", "lbm_last_error", "lbm_status_string",
    "lbm_abi_version", "lbm_device_count", "lbm_probe_pattern",
)


class LbmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


class lbm_params(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 6),
        ("precision", ctypes.c_int),
        ("inv_tau", ctypes.c_double),
        ("lid_velocity", ctypes.c_double),
        ("gravitation", ctypes.c_double * 3),
        ("store_densities", ctypes.c_int),
        ("store_velocities", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class lbm_physics(ctypes.Structure):
    _fields_ = [
        ("length_x", ctypes.c_double),
        ("size_x", ctypes.c_int),
        ("length_max", ctypes.c_double),
        ("velocity", ctypes.c_double * 3),
        ("acceleration", ctypes.c_double * 3),
        ("viscosity", ctypes.c_double),
        ("timestep", ctypes.c_double),
        ("max_velocity", ctypes.c_double),
        ("max_acceleration", ctypes.c_double),
    ]


class lbm_derived(ctypes.Structure):
    _fields_ = [
        ("velocity_dimless", ctypes.c_double * 3),
        ("acceleration_dimless", ctypes.c_double * 3),
        ("viscosity", ctypes.c_double),
        ("viscosity_dimless", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("inv_tau", ctypes.c_double),
        ("timestep", ctypes.c_double),
    ]


class lbm_batch_tuning(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int),
        ("cells_per_thread", ctypes.c_int),
        ("waves_per_cu", ctypes.c_int),
        ("options", ctypes.c_int * 2),
        ("pairs", ctypes.c_int),
        ("group_mb", ctypes.c_int),
    ]


class lbm_face_transfer(ctypes.Structure):
    _fields_ = [
        ("from_", ctypes.c_void_p),
        ("direction", ctypes.c_int),
        ("to", ctypes.c_void_p),
    ]


class lbm_device_arrays(ctypes.Structure):
    _fields_ = [
        ("dd", ctypes.c_void_p),
        ("flags", ctypes.c_void_p),
        ("densities", ctypes.c_void_p),
        ("velocities", ctypes.c_void_p),
        ("cells", ctypes.c_int64),
        ("plane_stride", ctypes.c_int64),
        ("size_h", ctypes.c_int * 3),
        ("precision", ctypes.c_int),
        ("element_bytes", ctypes.c_int),
    ]


_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP library {LIB_PATH} is not built; run `python -m cavityflow_cuda_amd.build` "
            "(hipcc --offload-arch=gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, V = ctypes.c_void_p, ctypes.c_int, None
    i3 = ctypes.POINTER(ctypes.c_int)
    sig = {
        "lbm_compute_parameters": ([ctypes.POINTER(lbm_physics), I, ctypes.POINTER(lbm_derived)], I),
        "lbm_create": ([ctypes.POINTER(lbm_params), ctypes.POINTER(P)], I),
        "lbm_destroy": ([P], I),
        "lbm_get_device_arrays": ([P, ctypes.POINTER(lbm_device_arrays)], I),
        "lbm_simulation_step_alpha": ([P, i3, i3, P], I),
        "lbm_simulation_step_beta": ([P, i3, i3, P], I),
        "lbm_get_density_distributions": ([P, i3, i3, P, I, P], I),
        "lbm_set_density_distributions": ([P, i3, i3, P, I, P], I),
        "lbm_set_density_distributions_dir": ([P, i3, i3, I, P, I, P], I),
        "lbm_get_flags": ([P, i3, i3, P], I),
        "lbm_set_flags": ([P, i3, i3, P], I),
        "lbm_get_velocities": ([P, i3, i3, P, I, P], I),
        "lbm_set_velocities": ([P, i3, i3, P, I, P], I),
        "lbm_get_densities": ([P, i3, i3, P, I, P], I),
        "lbm_set_densities": ([P, i3, i3, P, I, P], I),
        "lbm_face_box": ([P, I, I, i3, i3], I),
        "lbm_face_elements": ([P, I], ctypes.c_int64),
        "lbm_pack_face": ([P, I, I, P, P], I),
        "lbm_unpack_face": ([P, I, I, P, P], I),
        "lbm_unpack_face_pair": ([P, I, I, P, P, P], I),
        "lbm_pack_face_pair": ([P, I, I, P, P, P], I),
        "lbm_exchange_face": ([P, I, I, P, P], I),
        "lbm_exchange_faces": ([P, I, I, P], I),
        "lbm_device_alloc": ([I, ctypes.c_size_t, ctypes.POINTER(P)], I),
        "lbm_device_free": ([I, P], I),
        "lbm_set_launch_options": ([P, I, I], I),
        "lbm_get_launch_options": ([P, I, ctypes.POINTER(I)], I),
        "lbm_set_step_pack": ([P, P, P], I),
        "lbm_set_step_unpack": ([P, P, P], I),
        "lbm_get_moments": ([P, I, i3, i3, P, I, P], I),
        "lbm_batch_create": ([ctypes.POINTER(P), I, ctypes.POINTER(P)], I),
        "lbm_batch_destroy": ([P], I),
        "lbm_batch_size": ([P], I),
        "lbm_batch_set_tuning": ([P, ctypes.POINTER(lbm_batch_tuning)], I),
        "lbm_batch_get_tuning": ([P, ctypes.POINTER(lbm_batch_tuning)], I),
        "lbm_batch_step": ([P, I, P], I),
        "lbm_batch_run": ([P, I, I, P], I),
        "lbm_batch_launch": ([P, I, I, P], I),
        "lbm_batch_autotune": ([P, I, P, ctypes.POINTER(lbm_batch_tuning)], I),
        "lbm_batch_residual": ([P, I, ctypes.POINTER(ctypes.c_double), P], I),
        "lbm_reset": ([P], I),
        "lbm_state_bytes": ([P, ctypes.POINTER(ctypes.c_uint64)], I),
        "lbm_save_state": ([P, P, ctypes.c_uint64], I),
        "lbm_load_state": ([P, P, ctypes.c_uint64], I),
        "lbm_device_bytes": ([ctypes.POINTER(lbm_params), I, ctypes.POINTER(ctypes.c_uint64)], I),
        "lbm_reduce_box": ([P, i3, i3, ctypes.POINTER(ctypes.c_double), P], I),
        "lbm_reduce_boxes_noalloc": ([ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), I, P, P], I),
        "lbm_residual": ([P, I, ctypes.POINTER(ctypes.c_double), P], I),
        "lbm_sync": ([P, P], I),
        "lbm_last_error": ([], ctypes.c_char_p),
        "lbm_status_string": ([I], ctypes.c_char_p),
        "lbm_abi_version": ([], I),
        "lbm_device_count": ([ctypes.POINTER(I)], I),
        "lbm_probe_pattern": ([P, I, I, ctypes.POINTER(ctypes.c_uint64), P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int) -> int:
    if status != LBM_OK:
        L = lib()
        msg = L.lbm_last_error().decode() or L.lbm_status_string(status).decode()
        raise LbmError(status, msg)
    return status
