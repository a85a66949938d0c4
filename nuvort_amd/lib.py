"""ctypes binding of the C-ABI in include/nuvort.h (libnuvort_hip.so).

The library is required: loading fails loudly when it is missing or when no
GPU is visible to it; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnuvort_hip.so")

NV_OK = 0
ERRORS = {1: "NV_ERR_ARG", 2: "NV_ERR_HIP", 3: "NV_ERR_CAPACITY", 4: "NV_ERR_ZERO_LENGTH", 5: "NV_ERR_SINGULAR",
          6: "NV_ERR_NAN", 7: "NV_ERR_UNSUPPORTED", 8: "NV_ERR_TREE_DEPTH", 9: "NV_ERR_TIMESTEP", 10: "NV_ERR_STATE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NvParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("dt", C.c_double), ("quant_circ", C.c_double), ("corea", C.c_double),
                ("box_size", C.c_double * 3), ("boundary", C.c_int32 * 3), ("velocity", C.c_int32),
                ("tree_theta", C.c_double), ("tree_extra_correction", C.c_int32), ("normal_velocity", C.c_int32),
                ("norm_vel_xflow", C.c_double), ("alpha", C.c_double * 2), ("super_velocity", C.c_double * 3),
                ("normal_fluid_cutoff", C.c_double), ("t_zero_normal_fluid", C.c_double),
                ("norm_shear_omega", C.c_double), ("recon_shots", C.c_int32),
                ("overide_timestep_check", C.c_int32), ("precision", C.c_int32),
                ("near_radius", C.c_double), ("close_search", C.c_int32), ("exact_order", C.c_int32),
                ("capacity", C.c_int32), ("targets_per_lane", C.c_int32), ("kahan_block", C.c_int32),
                ("step_launch", C.c_int32), ("source_tile", C.c_int32)]


class NvState(C.Structure):
    _fields_ = [("pcount", C.c_int32), ("capacity", C.c_int32), ("itime", C.c_int32), ("recon_count", C.c_int32),
                ("remove_count", C.c_int32), ("wall_recon_count", C.c_int32), ("t", C.c_double),
                ("x", _dp), ("u", _dp), ("u1", _dp), ("u2", _dp), ("u_sup", _dp), ("u_mf", _dp),
                ("ghosti", _dp), ("ghostb", _dp), ("infront", _ip), ("behind", _ip), ("closest", _ip),
                ("closestd", _dp), ("pinnedi", _ip), ("pinnedb", _ip), ("wpinned", _ip),
                ("curv", _dp), ("stretch", _dp)]


class NvDiag(C.Structure):
    _fields_ = [("active", C.c_int32), ("recon_count", C.c_int32), ("remove_count", C.c_int32), ("itime", C.c_int32),
                ("t", C.c_double), ("maxu", C.c_double), ("maxdu", C.c_double), ("mean_u_sup", C.c_double),
                ("mean_u_mf", C.c_double), ("kappa_bar", C.c_double), ("kappa_min", C.c_double),
                ("kappa_max", C.c_double), ("total_length", C.c_double), ("avg_sep", C.c_double),
                ("tree_eval", C.c_int64), ("wall_recon_count", C.c_int32), ("pad", C.c_int32)]


class NvReconEvent(C.Structure):
    _fields_ = [("itime", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("tangle", C.c_int32),
                ("t", C.c_double), ("mid", C.c_double * 3), ("angle", C.c_double), ("dl", C.c_double)]


class NvTreeNode(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("width", C.c_double), ("cent", C.c_double * 3), ("circ", C.c_double * 3),
                ("pcount", C.c_int32), ("leaf_infront", C.c_int32)]


class NvStats(C.Structure):
    _fields_ = [("pair_evals", C.c_int64), ("near_evals", C.c_int64), ("tree_evals", C.c_int64),
                ("tree_nodes_visited", C.c_int64), ("tree_nodes", C.c_int32), ("tree_levels", C.c_int32),
                ("ms_velocity", C.c_double), ("ms_tree_build", C.c_double), ("ms_close", C.c_double),
                ("ms_topology", C.c_double), ("ms_step", C.c_double), ("point_updates", C.c_int64),
                ("pair_launches", C.c_int64), ("pair_ms_total", C.c_double), ("pair_evals_total", C.c_int64),
                ("walk_launches", C.c_int64), ("walk_ms_total", C.c_double), ("walk_nodes_total", C.c_int64),
                ("targets_per_lane", C.c_int32), ("kahan_block", C.c_int32), ("items_per_cu", C.c_int32),
                ("tuned", C.c_int32), ("tune_ms", C.c_double), ("tree_builds", C.c_int32),
                ("tree_sync_builds", C.c_int32), ("tree_replays", C.c_int32), ("tree_redone", C.c_int32),
                ("source_tile", C.c_int32), ("pad", C.c_int32)]


EXPORTS = ("nv_create", "nv_destroy", "nv_last_error", "nv_set_state", "nv_get_state", "nv_set_state_batch",
           "nv_get_state_batch", "nv_set_time", "nv_set_noise", "nv_ghostp",
           "nv_construct_tree", "nv_pmotion", "nv_velocity", "nv_pinsert", "nv_pclose", "nv_pclose_tree",
           "nv_precon", "nv_premove", "nv_wall_recon", "nv_enforce_boundary", "nv_empty_tree", "nv_nan_check", "nv_step",
           "nv_get_recon_events", "nv_get_tree", "nv_get_stats", "nv_autotune", "nv_pclose_noalloc",
           "nv_diagnostics", "nv_data_dump", "nv_data_restore", "nv_printf", "nv_print_info", "nv_tree_print",
           "nv_var_write", "nv_var_read", "nv_var_log_write")

_lib = None


class NuvortError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


def load(path: str = LIB_PATH):
    """Load the library and declare every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libnuvort_hip.so is not built (run __graft_entry__.build() or python -m nuvort_amd.build)")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "nv_create": [C.POINTER(NvParams), C.POINTER(vp)], "nv_destroy": [vp], "nv_set_state": [vp, C.POINTER(NvState)],
        "nv_set_noise": [vp, C.c_int32, C.POINTER(C.c_double)],
        "nv_get_state": [vp, C.POINTER(NvState)], "nv_set_time": [vp, C.c_double, C.c_int32],
        "nv_set_state_batch": [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(NvState)],
        "nv_get_state_batch": [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(NvState)],
        "nv_step": [vp, C.c_int32],
        "nv_get_recon_events": [vp, C.POINTER(NvReconEvent), C.c_int32, C.POINTER(C.c_int32)],
        "nv_get_tree": [vp, C.POINTER(NvTreeNode), C.c_int32, C.POINTER(C.c_int32)],
        "nv_get_stats": [vp, C.POINTER(NvStats)],
        "nv_autotune": [vp, C.c_char_p],
        # device pointers (x, infront, behind, closest, closestd) and a hipStream_t
        "nv_pclose_noalloc": [C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp],
        "nv_diagnostics": [vp, C.POINTER(NvDiag)], "nv_data_dump": [vp, C.c_char_p],
        "nv_data_restore": [vp, C.c_char_p], "nv_printf": [vp, C.c_char_p],
        "nv_print_info": [vp, C.POINTER(NvDiag), C.c_int32, C.c_char_p], "nv_tree_print": [vp, C.c_char_p],
        "nv_var_write": [C.c_char_p, C.POINTER(NvState)], "nv_var_read": [C.c_char_p, C.POINTER(NvState)],
        "nv_var_log_write": [C.c_char_p, C.POINTER(NvState)],
    }
    for name in EXPORTS:
        fn = getattr(L, name)
        fn.argtypes = sig.get(name, [vp])
        fn.restype = C.c_int
    L.nv_last_error.restype = C.c_char_p
    _lib = L
    return L
