"""ctypes binding of the C ABI declared in include/ikarus_amd.h (libikarus_amd.so).

Loading fails loudly when the HIP library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libikarus_amd.so")

IK_OK, IK_EINVAL, IK_EHIP, IK_ENEGDET, IK_ENOTIMPL, IK_ENOTBOUND, IK_ENOCONV, IK_ESINGULAR = 0, -1, -2, -3, -4, -5, -6, -7
DBC_RAW, DBC_FULL, DBC_REDUCED = 0, 1, 2
NEOHOOKE, SVK = 0, 1
FP64, FP32 = 0, 1
HOST, DEVICE = 0, 1
PCG_JACOBI_FP64, PCG_BJACOBI_FP32_IR, PCG_TWOLEVEL_FP32_IR = 0, 1, 2
PRED_CONSTANT, PRED_LINEAR, PRED_MLP = 0, 1, 2
PK2, KIRCHHOFF, CAUCHY = 0, 1, 2
VTK_POINT_DATA, VTK_CELL_DATA = 0, 1
KERNEL_ASSEMBLE_KR, KERNEL_ASSEMBLE_R, KERNEL_SPMV, KERNEL_MLP, KERNEL_GATHER_K = 0, 1, 2, 3, 4

EXPORTED = [
    "ik_create", "ik_destroy", "ik_last_error", "ik_bad_element", "ik_size", "ik_reduced_size",
    "ik_set_requirement", "ik_get_requirement", "ik_set_dirichlet_values", "ik_energy", "ik_stress_at", "ik_vector",
    "ik_matrix", "ik_assemble", "ik_matrix_csc", "ik_matrix_bsr",
    "ik_element", "ik_update_state", "ik_get_alpha", "ik_set_alpha", "ik_solve", "ik_newton_solve",
    "ik_load_control", "ik_path_following", "ik_mlp_num_params", "ik_mlp_set_params", "ik_mlp_get_params",
    "ik_mlp_set_load", "ik_mlp_predict",
    "ik_timing_enable", "ik_timing_read", "ik_timing_reset", "ik_synchronize", "ik_assembly_layout",
    "ik_mlp_create", "ik_mlp_destroy", "ik_mlp_last_error", "ik_mlp_params", "ik_mlp_load", "ik_mlp_forward",
    "ik_mlp_loss_grad", "ik_mlp_train_step", "ik_mlp_forward_noalloc", "ik_write_vtu",
]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("num_nodes", C.c_int64), ("num_elements", C.c_int64),
        ("X", C.c_void_p), ("conn", C.c_void_p), ("material", C.c_int32), ("lam", C.c_double),
        ("mu", C.c_double), ("eas_modes", C.c_int32), ("dirichlet", C.c_void_p), ("num_faces", C.c_int64),
        ("face_elem", C.c_void_p), ("face_axis", C.c_void_p), ("face_side", C.c_void_p),
        ("traction", C.c_double * 3), ("num_point_loads", C.c_int64), ("point_dof", C.c_void_p),
        ("point_coef", C.c_void_p),
    ]


class VtkField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("values", C.c_void_p), ("components", C.c_int32), ("cell", C.c_int32)]


class SolverOpts(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rtol", C.c_double), ("max_iter", C.c_int32), ("inner_rtol", C.c_double),
                ("max_refine", C.c_int32)]


class SolverInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("rel_residual", C.c_double),
                ("converged", C.c_int32)]


class NRSettings(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("min_iter", C.c_int32), ("precision", C.c_int32),
                ("linear", SolverOpts), ("forcing", C.c_double), ("eta_max", C.c_double)]


class NRInfo(C.Structure):
    _fields_ = [("success", C.c_int32), ("iterations", C.c_int32), ("residual_norm", C.c_double),
                ("correction_norm", C.c_double), ("linear_iterations", C.c_int64)]


class LCConfig(C.Structure):
    _fields_ = [("load_steps", C.c_int32), ("tbegin", C.c_double), ("tend", C.c_double), ("predictor", C.c_int32)]


class PFConfig(C.Structure):
    _fields_ = [("steps", C.c_int32), ("step_size", C.c_double), ("target_iterations", C.c_int32)]


class ControlInfo(C.Structure):
    _fields_ = [("success", C.c_int32), ("total_iterations", C.c_int32), ("linear_iterations", C.c_int64),
                ("solves", C.c_int32), ("fallbacks", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        L.ik_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(ProblemDesc)]
        L.ik_destroy.argtypes = [P]
        L.ik_destroy.restype = None
        L.ik_last_error.argtypes = [P]
        L.ik_last_error.restype = C.c_char_p
        L.ik_bad_element.argtypes = [P]
        L.ik_bad_element.restype = I64
        L.ik_size.argtypes = [P]
        L.ik_size.restype = I64
        L.ik_reduced_size.argtypes = [P]
        L.ik_reduced_size.restype = I64
        L.ik_set_requirement.argtypes = [P, P, D, I]
        L.ik_set_dirichlet_values.argtypes = [P, P, I]
        L.ik_energy.argtypes = [P, P]
        L.ik_stress_at.argtypes = [P, I, P, P, I]
        L.ik_get_requirement.argtypes = [P, P, P, I]
        L.ik_vector.argtypes = [P, I, P, I]
        L.ik_matrix.argtypes = [P, I, I]
        L.ik_assemble.argtypes = [P, I, I, P, I]
        L.ik_matrix_csc.argtypes = [P, I, P, P, P, P, P]
        L.ik_matrix_bsr.argtypes = [P, P, P, P, P, P, P, P]
        L.ik_element.argtypes = [P, I64, P, P]
        L.ik_update_state.argtypes = [P, P, I]
        L.ik_get_alpha.argtypes = [P, P, I]
        L.ik_set_alpha.argtypes = [P, P, I]
        L.ik_solve.argtypes = [P, P, P, I, C.POINTER(SolverOpts), C.POINTER(SolverInfo)]
        L.ik_newton_solve.argtypes = [P, C.POINTER(NRSettings), D, C.POINTER(NRInfo)]
        L.ik_load_control.argtypes = [P, C.POINTER(LCConfig), C.POINTER(NRSettings), C.POINTER(ControlInfo), P, P]
        L.ik_path_following.argtypes = [P, C.POINTER(PFConfig), C.POINTER(NRSettings), C.POINTER(ControlInfo), P, P,
                                        P]
        L.ik_mlp_num_params.argtypes = []
        L.ik_mlp_num_params.restype = I64
        L.ik_mlp_set_params.argtypes = [P, P, I]
        L.ik_mlp_get_params.argtypes = [P, P, I]
        L.ik_mlp_set_load.argtypes = [P, P, D]
        L.ik_mlp_predict.argtypes = [P, P, P, D, D, D, D, I]
        L.ik_timing_enable.argtypes = [P, I]
        L.ik_timing_read.argtypes = [P, I, P, P]
        L.ik_timing_reset.argtypes = [P]
        L.ik_synchronize.argtypes = [P]
        L.ik_assembly_layout.argtypes = [P, P, P, P]
        L.ik_mlp_create.argtypes = [C.POINTER(C.c_void_p), C.c_uint64]
        L.ik_mlp_destroy.argtypes = [P]
        L.ik_mlp_destroy.restype = None
        L.ik_mlp_last_error.argtypes = [P]
        L.ik_mlp_last_error.restype = C.c_char_p
        L.ik_mlp_params.argtypes = [P, P, I]
        L.ik_mlp_load.argtypes = [P, P, I]
        L.ik_mlp_forward.argtypes = [P, P, I64, P, I]
        L.ik_mlp_forward_noalloc.argtypes = [P, P, I64, P]
        L.ik_write_vtu.argtypes = [P, C.c_char_p, P, C.c_int32, C.c_int32, C.c_int32]
        L.ik_mlp_loss_grad.argtypes = [P, P, P, I64, P, P, I]
        L.ik_mlp_train_step.argtypes = [P, P, P, I64, C.c_float, P, I]
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class IkError(RuntimeError):
    def __init__(self, code, msg, bad_element=-1):
        super().__init__(f"ikarus_amd error {code}: {msg}")
        self.code = code
        self.bad_element = bad_element
