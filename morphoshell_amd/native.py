"""ctypes binding of the C-ABI in include/morphoshell_amd.h.

The shared library is built in-tree (``python -m morphoshell_amd.build``)
into ``morphoshell_amd/_lib/libmorphoshell_amd.so``.  Loading fails loudly if
it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmorphoshell_amd.so")

MS_OK, MS_ERR_ARG, MS_ERR_HIP, MS_ERR_PATCH_SEARCH, MS_ERR_MESH, MS_ERR_HIGH_FORCE, MS_ERR_STATE, MS_ERR_NO_DEVICE = range(8)
MS_FP64, MS_MIXED = 0, 1
MS_DIAL_BLEND, MS_DIAL_LCE = 1, 3

# Every symbol include/morphoshell_amd.h declares.
EXPORTED = [
    "ms_version", "ms_last_error", "ms_mesh_info", "ms_patch_fit", "ms_dof_masses", "ms_calc_scales",
    "ms_create", "ms_destroy", "ms_ctx_last_error", "ms_num_shells", "ms_set_state", "ms_get_state",
    "ms_set_dial_clock", "ms_get_dial_factor", "ms_set_shell_params", "ms_set_zero_dial_geometry",
    "ms_set_precision", "ms_get_precision", "ms_calc_deformation_forces", "ms_calc_non_deformation_forces",
    "ms_impose_constraints", "ms_advance_dynamics", "ms_is_at_equilibrium", "ms_calc_energies",
    "ms_calc_geometry", "ms_calc_geometry_noalloc", "ms_calc_stresses", "ms_calc_angle_deficits", "ms_calc_energy_densities", "ms_calc_diagnostics", "ms_get_forces", "ms_set_forces", "ms_step",
    "ms_steps_taken", "ms_resume_shell", "ms_set_tuning", "ms_get_tuning", "ms_enable_timing",
    "ms_last_step_timing", "ms_device_bytes", "ms_synchronize", "ms_calibration_copy", "ms_calibration_read",
    "ms_device_info", "ms_relax_setup", "ms_relax_run", "ms_set_check_log", "ms_read_check_log", "ms_nn_create", "ms_nn_destroy", "ms_nn_last_error", "ms_nn_num_params", "ms_nn_get_params",
    "ms_nn_set_params", "ms_nn_set_graph", "ms_nn_forward", "ms_nn_predict", "ms_nn_loss_backward", "ms_nn_adam_step",
    "ms_nn_gemm_test",
]


class MeshInfoOut(C.Structure):
    _fields_ = [
        ("perimeter", C.c_double), ("total_ref_area", C.c_double), ("approx_min_tri_size", C.c_double),
        ("num_edges", C.c_int32), ("num_boundary_edges", C.c_int32), ("num_boundary_nodes", C.c_int32),
        ("num_holes", C.c_int32),
    ]


class Scales(C.Structure):
    _fields_ = [
        ("poisson_ratio", C.c_double), ("ref_density", C.c_double), ("timestep_prefactor", C.c_double),
        ("damping_prefactor", C.c_double), ("dial_phase_time_prefactor", C.c_double),
        ("time_between_equil_checks_prefactor", C.c_double), ("outputs_per_char_long_time", C.c_double),
        ("gradient_descent", C.c_int32),
        ("min_ref_thickness", C.c_double), ("min_ref_shear_modulus", C.c_double),
        ("char_flexural_rigidity", C.c_double), ("char_force_scale", C.c_double), ("char_speed_scale", C.c_double),
        ("char_stretch_energy_density_scale", C.c_double), ("char_stretch_energy_scale", C.c_double),
        ("stretch_long_time", C.c_double), ("bend_long_time", C.c_double), ("char_long_time", C.c_double),
        ("damping_factor", C.c_double), ("timestep", C.c_double), ("dial_phase_time", C.c_double),
        ("timesteps_per_dial_phase", C.c_int64), ("timesteps_between_equil_checks", C.c_int64),
        ("steps_between_outputs", C.c_int64),
    ]


class ShellParams(C.Structure):
    _fields_ = [
        ("poisson_ratio", C.c_double), ("bend_stiffening_scale_factor", C.c_double), ("ref_density", C.c_double),
        ("min_ref_thickness", C.c_double), ("timestep", C.c_double), ("damping_factor", C.c_double),
        ("char_force_scale", C.c_double), ("char_speed_scale", C.c_double), ("gradient_descent", C.c_int32),
        ("dial_protocol", C.c_int32), ("dial_from_ansatz", C.c_int32), ("reserved", C.c_int32),
    ]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class NNArch(C.Structure):
    _fields_ = [
        ("node_features", C.c_int32), ("edge_features", C.c_int32), ("width", C.c_int32), ("layers", C.c_int32),
        ("out_dim", C.c_int32), ("reserved", C.c_int32), ("seed", C.c_uint64),
    ]


class NNGraphDesc(C.Structure):
    _fields_ = [
        ("num_graphs", C.c_int32), ("num_nodes", C.c_int32), ("num_edges", C.c_int32), ("reserved", C.c_int32),
        ("node_offsets", _ip), ("edge_src", _ip), ("edge_dst", _ip), ("node_feat", _fp), ("edge_feat", _fp),
    ]


class ShellDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_tris", C.c_int32), ("patch_ids", _ip), ("G", _dp), ("H", _dp),
        ("ref_pos", _dp), ("ref_areas", _dp), ("thickness", _dp), ("shear_modulus", _dp), ("abar_info", _dp),
        ("abar_info_cols", C.c_int32), ("bbar_info_cols", C.c_int32), ("bbar_info", _dp), ("a0", _dp), ("b0", _dp),
        ("node_masses", _dp), ("constrained", _u8p), ("params", ShellParams),
    ]


class StepReport(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("halted", C.c_int32), ("fail_step", C.c_int64), ("bad_dof", C.c_int64),
        ("bad_force", C.c_double), ("max_force_ratio", C.c_double), ("max_force_node", C.c_int32),
        ("max_speed_node", C.c_int32), ("max_speed_ratio", C.c_double), ("at_equilibrium", C.c_int32),
        ("checked", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("tri_block", C.c_int32), ("node_block", C.c_int32), ("tri_grid_per_cu", C.c_int32),
        ("node_grid_per_cu", C.c_int32), ("graph_steps", C.c_int32), ("fused", C.c_int32),
        ("tile_chunks", C.c_int32), ("tile_nodes", C.c_int32),
    ]


class RelaxParams(C.Structure):
    _fields_ = [
        ("tol_f", C.c_double), ("tol_v", C.c_double), ("equil_interval", C.c_int64),
        ("steps_per_dial_phase", C.c_int64), ("dial_spacing", C.c_double), ("dial_start", C.c_double),
        ("eta1", C.c_double), ("eta2", C.c_double), ("pause_on_eq", C.c_int32), ("reserved", C.c_int32),
    ]


class RelaxReport(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("halted", C.c_int32), ("failed", C.c_int32), ("num_eq", C.c_int32),
        ("eq_step", C.c_int64), ("fail_step", C.c_int64), ("dial_factor", C.c_double),
        ("max_force_ratio", C.c_double), ("max_speed_ratio", C.c_double), ("steps", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CheckRecord(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("dial_factor", C.c_double), ("max_force_ratio", C.c_double),
        ("max_speed_ratio", C.c_double), ("max_force_node", C.c_int32), ("max_speed_node", C.c_int32),
        ("at_equilibrium", C.c_int32), ("kind", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"native library {LIB_PATH} not found: build it with `python -m morphoshell_amd.build` "
            "(there is no fallback path)"
        )
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "ms_version": (C.c_char_p, []),
        "ms_last_error": (C.c_char_p, []),
        "ms_mesh_info": (C.c_int, [C.c_int32, C.c_int32, _dp, _ip, _dp, _dp, _u8p, C.POINTER(MeshInfoOut)]),
        "ms_patch_fit": (C.c_int, [C.c_int32, C.c_int32, _dp, _ip, C.c_double, _ip, _dp, _dp, _ip]),
        "ms_dof_masses": (C.c_int, [C.c_int32, C.c_int32, _ip, _dp, _dp, C.c_double, _dp]),
        "ms_calc_scales": (C.c_int, [C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double, C.POINTER(Scales)]),
        "ms_create": (C.c_int, [C.POINTER(ShellDesc), C.c_int32, C.c_int32, C.POINTER(P)]),
        "ms_destroy": (None, [P]),
        "ms_ctx_last_error": (C.c_char_p, [P]),
        "ms_num_shells": (C.c_int32, [P]),
        "ms_set_state": (C.c_int, [P, C.c_int32, _dp, _dp]),
        "ms_get_state": (C.c_int, [P, C.c_int32, _dp, _dp]),
        "ms_set_dial_clock": (C.c_int, [P, C.c_int32, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int32]),
        "ms_get_dial_factor": (C.c_int, [P, C.c_int32, _dp, C.POINTER(C.c_int64)]),
        "ms_set_shell_params": (C.c_int, [P, C.c_int32, C.POINTER(ShellParams)]),
        "ms_set_zero_dial_geometry": (C.c_int, [P, C.c_int32, _dp, _dp]),
        "ms_set_precision": (C.c_int, [P, C.c_int32]),
        "ms_get_precision": (C.c_int, [P]),
        "ms_calc_deformation_forces": (C.c_int, [P]),
        "ms_calc_non_deformation_forces": (C.c_int, [P]),
        "ms_impose_constraints": (C.c_int, [P]),
        "ms_advance_dynamics": (C.c_int, [P, C.POINTER(StepReport)]),
        "ms_is_at_equilibrium": (C.c_int, [P, C.c_int32, C.c_double, C.c_double, C.POINTER(StepReport)]),
        "ms_calc_energies": (C.c_int, [P, C.c_int32, _dp]),
        "ms_calc_geometry": (C.c_int, [P, C.c_int32, _dp, _dp, _dp]),
        "ms_calc_geometry_noalloc": (C.c_int, [P, C.c_int32, _dp, _dp, _dp]),
        "ms_calc_stresses": (C.c_int, [P, C.c_int32, _dp]),
        "ms_calc_angle_deficits": (C.c_int, [P, C.c_int32, C.c_void_p, _dp]),
        "ms_calc_energy_densities": (C.c_int, [P, C.c_int32, _dp]),
        "ms_calc_diagnostics": (C.c_int, [P, C.c_int32, _dp, _dp, _dp]),
        "ms_get_forces": (C.c_int, [P, C.c_int32, _dp]),
        "ms_set_forces": (C.c_int, [P, C.c_int32, _dp]),
        "ms_step": (C.c_int, [P, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int32, C.POINTER(StepReport)]),
        "ms_steps_taken": (C.c_int64, [P]),
        "ms_resume_shell": (C.c_int, [P, C.c_int32]),
        "ms_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "ms_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "ms_enable_timing": (C.c_int, [P, C.c_int32]),
        "ms_last_step_timing": (C.c_int, [P, _dp]),
        "ms_device_bytes": (C.c_int64, [P]),
        "ms_synchronize": (C.c_int, [P]),
        "ms_calibration_copy": (C.c_int, [C.c_int64, C.c_int32, _dp]),
        "ms_calibration_read": (C.c_int, [C.c_int64, C.c_int32, _dp]),
        "ms_device_info": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "ms_relax_setup": (C.c_int, [P, C.c_int32, C.POINTER(RelaxParams)]),
        "ms_relax_run": (C.c_int, [P, C.c_int64, C.c_int64, C.POINTER(RelaxReport)]),
        "ms_set_check_log": (C.c_int, [P, C.c_int32]),
        "ms_read_check_log": (C.c_int, [P, C.c_int32, C.POINTER(CheckRecord), C.c_int32,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
        "ms_nn_create": (C.c_int, [C.POINTER(NNArch), C.POINTER(P)]),
        "ms_nn_destroy": (None, [P]),
        "ms_nn_last_error": (C.c_char_p, [P]),
        "ms_nn_num_params": (C.c_int64, [P]),
        "ms_nn_get_params": (C.c_int, [P, _fp]),
        "ms_nn_set_params": (C.c_int, [P, _fp]),
        "ms_nn_set_graph": (C.c_int, [P, C.POINTER(NNGraphDesc)]),
        "ms_nn_forward": (C.c_int, [P, _fp]),
        "ms_nn_predict": (C.c_int, [P, C.POINTER(NNGraphDesc), _fp]),
        "ms_nn_loss_backward": (C.c_int, [P, _fp, _fp, _fp]),
        "ms_nn_adam_step": (C.c_int, [P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
        "ms_nn_gemm_test": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _fp, _fp, _fp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code: int, ctx=None) -> None:
    if code != MS_OK:
        L = lib()
        msg = (L.ms_ctx_last_error(ctx) if ctx else L.ms_last_error()) or b""
        raise NativeError(code, msg.decode())


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def u8ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)
