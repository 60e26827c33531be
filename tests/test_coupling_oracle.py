"""The coupling oracle (oracle/coupling.py) pinned by closed forms: planar
angles and rotation matrices of known quaternions, the position mapping on
representable numbers, and the load conversions of each mode."""
import math

import numpy as np
import pytest

from oracle import coupling as O

F = np.float32


@pytest.mark.parametrize("theta", [0.0, 0.3, -1.2, 2.9])
def test_planar_angle_and_matrix_of_z_rotation(theta):
    q = (math.cos(theta / 2), 0.0, 0.0, math.sin(theta / 2))
    assert abs(float(O.planar_angle(q)) - theta) < 2e-6
    R = O.quat_matrix(q).reshape(3, 3)
    c, s = math.cos(theta), math.sin(theta)
    np.testing.assert_allclose(R, [[c, -s, 0], [s, c, 0], [0, 0, 1]], atol=2e-7)


def test_matrix_of_quarter_turn_about_x_and_normalisation():
    R = O.quat_matrix((2.0, 2.0, 0.0, 0.0)).reshape(3, 3)  # unnormalised (cos 45, sin 45, 0, 0) * 2 sqrt 2
    np.testing.assert_allclose(R, [[1, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-7)
    # a rotation matrix: orthonormal, determinant 1
    q = np.array([0.3, -0.5, 0.7, 0.2])
    R = O.quat_matrix(q).reshape(3, 3).astype(np.float64)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert abs(np.linalg.det(R) - 1) < 1e-6


def test_position_mapping_exact():
    assert O.lattice_position(1.5, 0.25, 10.0, 8.0) == F(20.0)
    assert O.lattice_position(-0.5, 0.5, 64.0, 16.0) == F(48.0)


def test_load_modes():
    row = [8.0, -4.0, 2.0, 1.0, -3.0, 6.0]
    Fv, T = O.engine_loads(row, 9, 0, s=2.0)
    assert list(Fv) == [4.0, -2.0, 0.0] and list(T) == [0.0, 0.0, 3.0]     # 2D: torque / scale once
    Fv, T = O.engine_loads(row, 27, 0, s=2.0)
    assert list(Fv) == [4.0, -2.0, 1.0] and list(T) == [0.25, -0.75, 1.5]  # 3D: torque / scale^2
    Fv, T = O.engine_loads(row, 27, 1, fconv=0.5, tconv=0.25)
    assert list(Fv) == [4.0, -2.0, 1.0] and list(T) == [0.25, -0.75, 1.5]
