"""CPU restatement of the coupling kernels (test infrastructure only).

FP32 in the kernels' operation order:
  pose conversion   convert_and_update_solid_batch_2d  envs/lbm/lbm_fluid_env_func.py:72-105
                    convert_and_update_solid_batch_3d  envs/lbm3d/lbm_fluid_env_3d_func.py:42-99
  load conversion   extract_forces_torques_batch       envs/lbm/lbm_fluid_env_func.py:51-69
                    extract_forces_torques_batch_3d / _physical_3d
                                                       envs/lbm3d/lbm_fluid_env_3d_func.py:102-185
The quaternion -> matrix map and the free-body step have no reference
counterpart (the reference hands poses to Warp transforms and steps
MuJoCo-Warp); they are pinned by closed forms in the tests.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def lattice_position(p, eo, lo, s):
    """(engine - engine_origin) * scale + lattice_origin, FP32."""
    return (F(p) - F(eo)) * F(s) + F(lo)


def planar_angle(q):
    """atan2(2 (w z + x y), 1 - 2 (y^2 + z^2)) of a (w, x, y, z) quaternion."""
    w, x, y, z = (F(v) for v in q)
    return np.arctan2(F(2.0) * (w * z + x * y), F(1.0) - F(2.0) * (y * y + z * z))


def quat_matrix(q):
    """Row-major rotation of a (w, x, y, z) quaternion, normalised first, FP32."""
    w, x, y, z = (F(v) for v in q)
    n = np.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    one, two = F(1.0), F(2.0)
    return np.array([one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y),
                     two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x),
                     two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)], F)


def engine_loads(row, lattice, mode, s=1.0, fconv=1.0, tconv=1.0):
    """(force, torque) FP32 of one FP64 row F(3), tau(3)."""
    Fv = np.array(row[:3], np.float64).astype(F)
    T = np.array(row[3:6], np.float64).astype(F)
    if lattice == 9:
        Fv[2] = 0
        T[0] = T[1] = 0
    if mode == 0:
        s = F(s)
        ts = s if lattice == 9 else s * s
        Fv, T = Fv / s, T / ts
        if lattice == 9:
            Fv[2] = 0
            T[0] = T[1] = 0
    else:
        Fv, T = Fv * F(fconv), T * F(tconv)
    return Fv, T
