"""ctypes binding of libfem_hip.so (C-ABI declared in include/fem_hip.h).

The library is the only compute path: if it is missing or cannot be loaded
the import fails with an error instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfem_hip.so"

FEM_OK = 0
FEM_ERR_ARG, FEM_ERR_HIP, FEM_ERR_UNSUPPORTED, FEM_ERR_STATE = 1, 2, 3, 4
FEM_TERM_TET, FEM_TERM_HEX, FEM_TERM_GROUND, FEM_TERM_GRAVITY, FEM_TERM_DAMPING = 1, 2, 4, 8, 16
FEM_TERM_SPRING, FEM_TERM_TRI, FEM_TERM_PARTICLE, FEM_TERM_SOFT = 32, 64, 128, 256

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("particle_count", ctypes.c_int32),
        ("particle_q", _fp),
        ("particle_mass", _fp),
        ("particle_inv_mass", _fp),
        ("particle_radius", _fp),
        ("particle_flags", _ip),
        ("tet_count", ctypes.c_int32),
        ("tet_indices", _ip),
        ("tet_poses", _fp),
        ("tet_materials", _fp),
        ("hex_count", ctypes.c_int32),
        ("hex_indices", _ip),
        ("hex_inv_J0", _fp),
        ("hex_det_J0_w", _fp),
        ("hex_materials", _fp),
        ("fixed_count", ctypes.c_int32),
        ("fixed_indices", _ip),
        ("instance_count", ctypes.c_int32),
        ("instance_offsets", _ip),
        ("spring_count", ctypes.c_int32),
        ("spring_indices", _ip),
        ("spring_rest_length", _fp),
        ("spring_stiffness", _fp),
        ("spring_damping", _fp),
        ("tri_count", ctypes.c_int32),
        ("tri_indices", _ip),
        ("tri_poses", _fp),
        ("tri_materials", _fp),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("min_stretch", ctypes.c_float),
        ("max_stretch", ctypes.c_float),
        ("has_ground", ctypes.c_int32),
        ("ground", ctypes.c_float * 4),
        ("ground_ke", ctypes.c_float),
        ("ground_kd", ctypes.c_float),
        ("ground_kf", ctypes.c_float),
        ("ground_mu", ctypes.c_float),
        ("linear_damping", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("enable_tet_activations", ctypes.c_int32),
        ("tolerance", ctypes.c_float),
        ("maxiter", ctypes.c_int32),
        ("early_exit", ctypes.c_int32),
    ]


# Every exported symbol with (restype, argtypes).
SIGNATURES = {
    "fem_create": (ctypes.c_int, [ctypes.POINTER(ModelDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "fem_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fem_last_error": (ctypes.c_char_p, []),
    "fem_pattern_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _ip]),
    "fem_memory_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "fem_bind": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [ctypes.c_int64]),
    "fem_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Params)]),
    "fem_set_particle_contact": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_float] * 5),
    "fem_set_soft_contacts": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_int] + [ctypes.c_float] * 4),
    "fem_set_solver": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fem_set_spring_torque": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_float, ctypes.c_float]),
    "fem_bind_gravity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fem_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fem_get_launch": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip]),
    "fem_set_cg_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fem_restore_fixed": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fem_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]),
    "fem_step": (ctypes.c_int, [ctypes.c_void_p] * 8 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_step_rhs": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_integrate_rhs": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_step_solve": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_step_solve_static": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_eval_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 5),
    "fem_solve_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _ip, _fp, _fp, _ip]),
    "fem_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_void_p]),
    "fem_vjp": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6 + [ctypes.c_float] + [ctypes.c_void_p] * 5),
    "fem_set_tet_materials": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fem_apply_operator": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fem_apply_preconditioner": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fem_set_barrier_spin_limit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "fem_block_jacobi_mv": (ctypes.c_int, [ctypes.c_void_p] * 4 + [ctypes.c_float, ctypes.c_float, ctypes.c_void_p]),
    "fem_diagonal_blocks": (ctypes.c_int, [ctypes.c_void_p] * 3),
    "fem_add_load": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "fem_solve_increment": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "fem_reconstruct": (ctypes.c_int, [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_void_p]),
    "fem_diff_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_int32]),
    "fem_prepare_topology": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fem_update_topology": (ctypes.c_int, [ctypes.c_void_p] * 9 + [ctypes.c_float, _ip, ctypes.c_void_p]),
    "fem_bind_component_scaling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 5),
    "fem_refresh_component_scaling": (ctypes.c_int, [ctypes.c_void_p] * 4),
}

_lib = None


class FEMError(RuntimeError):
    pass


def load():
    """Load libfem_hip.so once; raises if it is absent (no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python newton-solver-fem_amd/build.py` "
                "(the HIP library is the only compute path)"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = ""):
    if rc != FEM_OK:
        msg = load().fem_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(f"{what}: {msg}")
        raise FEMError(f"{what} failed ({rc}): {msg}")
