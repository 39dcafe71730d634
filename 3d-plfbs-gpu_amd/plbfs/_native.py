"""ctypes binding of libplbfs.so (include/plbfs_hip.h, include/plbfs_case.h).

The library is built in-tree (3d-plfbs-gpu_amd/build/libplbfs.so).  Loading
it never falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "build", "libplbfs.so")

OK = 0
ERR_NAMES = {0: "ok", -1: "bad input", -2: "nan", -3: "rho blow-up", -4: "unsupported bc type",
             -5: "hip error", -6: "no device", -7: "ibm node left support"}

FP64, FP32 = 0, 1
GRAD_LS, GRAD_HYBRID, GRAD_GG = 0, 1, 2
DT_LOCAL, DT_CONSTANT, DT_MIN, DT_TALTS = 0, 1, 2, 3
DELTA_HAT2, DELTA_PESKIN4 = 0, 1

# every symbol the two headers declare
EXPORTS = [
    "plbfs_last_error", "plbfs_device_count", "plbfs_create", "plbfs_destroy", "plbfs_set_state",
    "plbfs_set_ghost_state", "plbfs_get_state", "plbfs_set_ibm", "plbfs_get_ibm", "plbfs_get_body_force", "plbfs_run",
    "plbfs_run_noalloc", "plbfs_get_residual_history", "plbfs_eval_residual", "plbfs_get_dt", "plbfs_enable_timing",
    "plbfs_kernel_stats", "plbfs_layout_info", "plbfs_wall_forces", "plbfs_ibm_body_force",
    "plbfs_mesh_read_openfoam", "plbfs_mesh_from_arrays", "plbfs_mesh_free", "plbfs_mesh_info",
    "plbfs_mesh_view_get", "plbfs_mesh_bface_patch", "plbfs_mesh_patch_type", "plbfs_mesh_raw",
    "plbfs_mesh_n_face_verts", "plbfs_case_load", "plbfs_case_free", "plbfs_case_params",
    "plbfs_case_string", "plbfs_case_number", "plbfs_case_patches", "plbfs_case_initial_state",
    "plbfs_case_n_objects", "plbfs_case_object_nodes", "plbfs_write_tecplot", "plbfs_write_tecplot_grid",
    "plbfs_write_tecplot_nodes", "plbfs_write_fields",
    "plbfs_read_fields", "plbfs_program_run",
    "plbfs_batch_create", "plbfs_batch_destroy", "plbfs_batch_set_state", "plbfs_batch_get_state",
    "plbfs_batch_set_ibm", "plbfs_batch_run", "plbfs_batch_set_tiling", "plbfs_batch_autotune",
    "plbfs_batch_layout_info",
    "plbfs_mpnn_n_params", "plbfs_mpnn_create", "plbfs_mpnn_destroy", "plbfs_mpnn_set_params",
    "plbfs_mpnn_get_params", "plbfs_mpnn_forward", "plbfs_mpnn_loss_grad", "plbfs_mpnn_adam_step",
    "plbfs_mpnn_train",
]


class MeshView(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
                ("owner", C.POINTER(C.c_int32)), ("neighbour", C.POINTER(C.c_int32)),
                ("cell_centroid", C.POINTER(C.c_double)), ("cell_volume", C.POINTER(C.c_double)),
                ("face_centroid", C.POINTER(C.c_double)), ("face_normal", C.POINTER(C.c_double)),
                ("face_area", C.POINTER(C.c_double)), ("face_dt", C.POINTER(C.c_double))]


class Patch(C.Structure):
    _fields_ = [("rho_type", C.c_int32), ("vel_type", C.c_int32), ("rho", C.c_double), ("u", C.c_double),
                ("v", C.c_double), ("w", C.c_double), ("is_empty", C.c_int32), ("per_translate", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [("visc", C.c_double), ("gamma", C.c_double), ("fneq_min", C.c_double), ("cfl", C.c_double),
                ("max_velocity", C.c_double), ("ref_rho", C.c_double), ("force", C.c_double),
                ("tolerance", C.c_double), ("gradient_type", C.c_int32), ("time_stepping", C.c_int32),
                ("precision", C.c_int32), ("check_interval", C.c_int32), ("lsq_exact_det", C.c_int32),
                ("domain_x", C.c_double), ("domain_y", C.c_double)]


class IbmNodes(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("x_ref", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
                ("stiffness", C.c_double), ("node_area", C.c_double), ("dt", C.c_double), ("h", C.c_double),
                ("kernel", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_int32), ("converged", C.c_int32),
                ("total_steps", C.c_int32), ("residual", C.c_double * 4), ("solve_ms", C.c_double)]


_lib = None


class PlbfsError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__("plbfs error %d (%s): %s" % (rc, ERR_NAMES.get(rc, "?"), msg))
        self.rc = rc


def lib():
    """Load libplbfs.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libplbfs.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.plbfs_last_error.restype = C.c_char_p
    L.plbfs_device_count.restype = C.c_int
    L.plbfs_create.argtypes = [C.POINTER(MeshView), C.POINTER(Params), C.POINTER(Patch), C.c_int32, ip,
                               C.POINTER(vp)]
    L.plbfs_destroy.argtypes = [vp]
    L.plbfs_set_state.argtypes = [vp, dp, dp, dp, dp]
    L.plbfs_get_state.argtypes = [vp, dp, dp, dp, dp]
    L.plbfs_set_ghost_state.argtypes = [vp, dp, dp, dp, dp]
    L.plbfs_set_ibm.argtypes = [vp, C.POINTER(IbmNodes)]
    L.plbfs_get_ibm.argtypes = [vp, dp, dp, dp]
    L.plbfs_get_body_force.argtypes = [vp, dp]
    L.plbfs_run.argtypes = [vp, C.c_int32, C.POINTER(Result)]
    L.plbfs_run_noalloc.argtypes = [vp, C.c_int32, C.POINTER(Result), dp]
    L.plbfs_get_residual_history.argtypes = [vp, dp, C.c_int32]
    L.plbfs_eval_residual.argtypes = [vp, dp, dp]
    L.plbfs_get_dt.argtypes = [vp, dp]
    L.plbfs_enable_timing.argtypes = [vp, C.c_int32]
    L.plbfs_kernel_stats.argtypes = [vp, C.c_int32, dp, ip]
    L.plbfs_layout_info.argtypes = [vp, i64p]
    L.plbfs_wall_forces.argtypes = [vp, C.POINTER(MeshView), ip, C.c_int32, dp]
    L.plbfs_ibm_body_force.argtypes = [vp, dp]
    L.plbfs_mesh_read_openfoam.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.POINTER(vp)]
    L.plbfs_mesh_from_arrays.argtypes = [C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_int32, ip, ip, ip, ip,
                                         C.c_int32, C.POINTER(C.c_char_p), ip, C.c_double, C.c_double, C.c_double,
                                         C.POINTER(vp)]
    L.plbfs_mesh_free.argtypes = [vp]
    L.plbfs_mesh_info.argtypes = [vp, i64p]
    L.plbfs_mesh_view_get.argtypes = [vp, C.POINTER(MeshView)]
    L.plbfs_mesh_bface_patch.argtypes = [vp]
    L.plbfs_mesh_bface_patch.restype = ip
    L.plbfs_mesh_patch_type.argtypes = [vp, C.c_int32, C.c_char_p, C.c_int32]
    L.plbfs_mesh_raw.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), ip, ip,
                                 ip, ip]
    L.plbfs_mesh_n_face_verts.argtypes = [vp]
    L.plbfs_mesh_n_face_verts.restype = C.c_int64
    L.plbfs_batch_create.argtypes = [C.POINTER(MeshView), C.POINTER(Params), C.POINTER(Patch), C.c_int32, ip,
                                     C.c_int32, dp, C.POINTER(vp)]
    L.plbfs_batch_destroy.argtypes = [vp]
    L.plbfs_batch_set_state.argtypes = [vp, C.c_int32, dp, dp, dp, dp]
    L.plbfs_batch_get_state.argtypes = [vp, C.c_int32, dp, dp, dp, dp]
    L.plbfs_batch_set_ibm.argtypes = [vp, C.c_int32, C.POINTER(IbmNodes)]
    L.plbfs_batch_run.argtypes = [vp, C.c_int32, C.POINTER(Result)]
    L.plbfs_batch_set_tiling.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.plbfs_batch_autotune.argtypes = [vp, C.c_int32, ip, ip, ip, dp]
    L.plbfs_batch_layout_info.argtypes = [vp, i64p]
    fp = C.POINTER(C.c_float)
    L.plbfs_mpnn_n_params.restype = C.c_int
    L.plbfs_mpnn_create.argtypes = [C.c_int32, C.c_int32, ip, ip, fp, C.c_int32, C.POINTER(vp)]
    L.plbfs_mpnn_destroy.argtypes = [vp]
    L.plbfs_mpnn_set_params.argtypes = [vp, fp]
    L.plbfs_mpnn_get_params.argtypes = [vp, fp]
    L.plbfs_mpnn_forward.argtypes = [vp, fp, fp]
    L.plbfs_mpnn_loss_grad.argtypes = [vp, fp, fp, fp, dp, fp]
    L.plbfs_mpnn_adam_step.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.plbfs_mpnn_train.argtypes = [vp, C.c_int32, fp, fp, fp, C.c_int32, C.c_float, C.c_uint64, dp]
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise PlbfsError(rc, lib().plbfs_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))
