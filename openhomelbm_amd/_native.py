"""ctypes binding of libhomelbm.so (the C ABI in include/homelbm.h).

There is no fallback: if the library is missing or a call fails, this module
raises. The library is built in-tree by openhomelbm_amd.build_native.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build_native

HL_D2Q9, HL_D3Q27 = 9, 27
HL_F32, HL_F64, HL_F16, HL_BF16 = 0, 1, 2, 3
HL_KERNEL_AUTO, HL_KERNEL_GATHER, HL_KERNEL_TILED = 0, 1, 2
HL_MAX_SOLIDS = 64

#: every entry point declared in include/homelbm.h
EXPORTS = ("hl_version", "hl_env_bytes", "hl_last_error", "hl_step", "hl_forces", "hl_init_uniform",
           "hl_voxelize", "hl_voxelize_polygons", "hl_mass", "hl_set_tiling", "hl_get_tiling",
           "hl_batch_create", "hl_batch_destroy", "hl_batch_set_bodies", "hl_batch_set_state", "hl_batch_set_strips",
           "hl_batch_get_state", "hl_batch_step", "hl_batch_capture", "hl_batch_replay", "hl_batch_size",
           "hl_render_plane", "hl_render_projection", "hl_couple_poses", "hl_couple_forces",
           "hl_free_bodies_step", "hl_body_capsules", "hl_voxelize_mesh", "hl_step_polygons",
           "hl_step_polygons_worlds", "hl_kernel_timer", "hl_kernel_timer_read", "hl_copy_state",
           "hl_step_noalloc", "hl_step_capsules", "hl_batch_run", "hl_batch_run_status",
           "hl_chain_create", "hl_chain_destroy", "hl_chain_step", "hl_step_worlds")
HL_COUPLE_SCALED, HL_COUPLE_PHYSICAL = 0, 1
# render kinds (include/homelbm.h)
HL_RENDER_SPEED, HL_RENDER_VORTICITY, HL_RENDER_VORTICITY_SOLID, HL_RENDER_RGB, HL_RENDER_SOLID = 0, 1, 2, 3, 4
HL_PROJECT_FIRST_SPEED, HL_PROJECT_MAX_SPEED, HL_PROJECT_MAX_VORTICITY = 0, 1, 2


HL_CHAIN_MAX_BODIES = HL_CHAIN_MAX_DOFS = HL_CHAIN_MAX_ACTS = 16
HL_CHAIN_SLIDE, HL_CHAIN_HINGE = 0, 1
HL_CHAIN_RK4, HL_CHAIN_EULER = 0, 1


class HLChainModel(C.Structure):
    """Mirror of struct hl_chain_model."""

    _NB, _NV, _NA = HL_CHAIN_MAX_BODIES, HL_CHAIN_MAX_DOFS, HL_CHAIN_MAX_ACTS
    _fields_ = [
        ("nbody", C.c_int32), ("ndof", C.c_int32), ("nact", C.c_int32), ("integrator", C.c_int32),
        ("parent", C.c_int32 * _NB), ("free_body", C.c_int32 * _NB),
        ("dof_body", C.c_int32 * _NV), ("dof_type", C.c_int32 * _NV),
        ("act_dof", C.c_int32 * _NA), ("ctrl_limited", C.c_int32 * _NA),
        ("pos", (C.c_double * 2) * _NB), ("angle", C.c_double * _NB), ("com", (C.c_double * 2) * _NB),
        ("mass", C.c_double * _NB), ("inertia", C.c_double * _NB), ("posz", C.c_double * _NB),
        ("comz", C.c_double * _NB),
        ("dof_axis", (C.c_double * 2) * _NV), ("dof_anchor", (C.c_double * 2) * _NV), ("dof_sign", C.c_double * _NV),
        ("damping", C.c_double * _NV), ("stiffness", C.c_double * _NV), ("springref", C.c_double * _NV),
        ("armature", C.c_double * _NV),
        ("gear", C.c_double * _NA), ("ctrl_lo", C.c_double * _NA), ("ctrl_hi", C.c_double * _NA),
        ("gravity", C.c_double * 2), ("dt", C.c_double), ("applied_factor", C.c_double),
    ]


class HLEnv(C.Structure):
    """Mirror of struct hl_env."""

    _fields_ = [
        ("lattice", C.c_int32), ("dtype", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("nsolid", C.c_int32),
        ("stride", C.c_int64),
        ("moments", C.c_void_p * 2),
        ("mask", C.c_void_p),
        ("mask_box", C.c_int32 * 6),
        ("force", C.c_void_p),
        ("solids", C.c_void_p),
        ("nu", C.c_float),
        ("bc_type", C.c_int32 * 6),
        ("bc_value", (C.c_float * 3) * 6),
        ("scale", C.c_float * 10),
        ("offset", C.c_float * 10),
        ("mask_epoch", C.c_uint64),
        ("rpitch", C.c_int64),
        ("force_stride", C.c_int64),
    ]


class HLError(RuntimeError):
    pass


_lib = None


def lib_path() -> str:
    return build_native.LIB


def load():
    """Load libhomelbm.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise HLError(f"{path} is missing: build it with `python -m openhomelbm_amd.build_native` "
                      "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(path)
    P, I32 = C.c_void_p, C.c_int32
    L.hl_version.restype = I32
    L.hl_env_bytes.restype = I32
    L.hl_last_error.restype = C.c_char_p
    env_p = C.POINTER(HLEnv)
    L.hl_step.argtypes = [env_p, I32, I32, I32, P, P]
    L.hl_step_noalloc.argtypes = [env_p, I32, I32, I32, P, P]
    L.hl_forces.argtypes = [env_p, I32, P, P]
    L.hl_init_uniform.argtypes = [env_p, I32, C.c_float, C.c_float, C.c_float, P]
    L.hl_voxelize.argtypes = [env_p, P, P, I32, P, P]
    L.hl_voxelize_polygons.argtypes = [env_p, P, P, P, I32, I32, P, P]
    L.hl_mass.argtypes = [env_p, I32, P, P]
    L.hl_set_tiling.argtypes = [I32, I32, I32, I32]
    L.hl_get_tiling.argtypes = [I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]
    PI32 = C.POINTER(I32)
    L.hl_batch_create.argtypes = [env_p, I32, C.POINTER(C.c_void_p)]
    L.hl_batch_destroy.argtypes = [P]
    L.hl_batch_destroy.restype = None
    L.hl_batch_set_bodies.argtypes = [P, I32, P, PI32, P, I32, P, I32]
    L.hl_batch_set_state.argtypes = [P, PI32, PI32]
    L.hl_batch_set_strips.argtypes = [P, I32, I32]
    L.hl_batch_get_state.argtypes = [P, PI32, PI32]
    L.hl_batch_step.argtypes = [P, P, I32, P, I32, P]
    L.hl_batch_capture.argtypes = [P, P, I32, P]
    L.hl_batch_replay.argtypes = [P, P]
    L.hl_batch_size.argtypes = [P, PI32, PI32, PI32]
    L.hl_batch_run.argtypes = [P, P, I32, P, I32, P]
    L.hl_batch_run_status.argtypes = [P, PI32]
    F32 = C.c_float
    L.hl_render_plane.argtypes = [env_p, I32, I32, I32, I32, F32, F32, P, P]
    L.hl_render_projection.argtypes = [env_p, I32, I32, I32, P, P]
    L.hl_couple_poses.argtypes = [I32, I32, I32, I32, P, I32, P, I32, P, P, P, P, P, P, P]
    L.hl_couple_forces.argtypes = [I32, I32, I32, I32, I32, P, C.c_int64, P, P, F32, F32, P, P, P, I32, P, P]
    L.hl_free_bodies_step.argtypes = [I32, I32, P, P, P, P, F32, F32, F32, F32, P]
    L.hl_body_capsules.argtypes = [I32, I32, P, I32, P, I32, P, P, P, P, P, P, P]
    L.hl_voxelize_mesh.argtypes = [env_p, P, P, P, I32, P, I32, P, P]
    L.hl_step_polygons.argtypes = [env_p, I32, I32, P, P, P, I32, P, P, P]
    L.hl_step_capsules.argtypes = [env_p, P, I32, P, I32, I32, I32, P, P]
    L.hl_step_polygons_worlds.argtypes = [env_p, I32, I32, I32, P, P, P, P, P, P, P]
    L.hl_step_worlds.argtypes = [env_p, I32, I32, I32, I32, P, P]
    L.hl_kernel_timer.argtypes = [I32, I32]
    L.hl_kernel_timer_read.argtypes = [C.POINTER(C.c_double), PI32]
    L.hl_copy_state.argtypes = [env_p, I32, P]
    L.hl_chain_create.argtypes = [C.POINTER(HLChainModel), C.POINTER(C.c_void_p)]
    L.hl_chain_destroy.argtypes = [P]
    L.hl_chain_step.argtypes = [P, I32, P, P, P, P, I32, P, P, P, P, P]
    for name in ("hl_step", "hl_forces", "hl_init_uniform", "hl_voxelize", "hl_voxelize_polygons", "hl_mass",
                 "hl_set_tiling", "hl_get_tiling", "hl_batch_create", "hl_batch_set_bodies", "hl_batch_set_state",
                 "hl_batch_set_strips", "hl_batch_get_state", "hl_batch_step", "hl_batch_capture", "hl_batch_replay", "hl_batch_size",
                 "hl_render_plane", "hl_render_projection", "hl_couple_poses", "hl_couple_forces",
                 "hl_free_bodies_step", "hl_body_capsules", "hl_voxelize_mesh", "hl_step_polygons",
                 "hl_step_polygons_worlds", "hl_kernel_timer", "hl_kernel_timer_read", "hl_copy_state",
                 "hl_step_noalloc", "hl_step_capsules", "hl_batch_run", "hl_batch_run_status",
                 "hl_chain_create", "hl_chain_destroy", "hl_chain_step", "hl_step_worlds"):
        getattr(L, name).restype = I32
    if L.hl_env_bytes() != C.sizeof(HLEnv):
        raise HLError(f"hl_env layout mismatch: library {L.hl_env_bytes()} bytes, binding {C.sizeof(HLEnv)}")
    _lib = L
    return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().hl_last_error().decode(errors="replace")
        raise HLError(f"{what} failed ({rc}): {msg}")


def set_tiling(lattice: int, tx: int, ty: int, zc: int) -> None:
    check(load().hl_set_tiling(lattice, tx, ty, zc), "hl_set_tiling")


def kernel_timer(max_launches: int, every: int = 1) -> None:
    """Record HIP events around one step-kernel launch in every `every`, up to
    max_launches recorded launches (0 stops)."""
    check(load().hl_kernel_timer(int(max_launches), int(every)), "hl_kernel_timer")


def kernel_timer_read():
    """(summed step-kernel launch time in ms, launches) since the timer was
    armed or last read; waits for the recorded launches."""
    ms, n = C.c_double(), C.c_int32()
    check(load().hl_kernel_timer_read(C.byref(ms), C.byref(n)), "hl_kernel_timer_read")
    return ms.value, n.value


def get_tiling(lattice: int):
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    check(load().hl_get_tiling(lattice, C.byref(a), C.byref(b), C.byref(c)), "hl_get_tiling")
    return a.value, b.value, c.value
