"""GPU coupling kernels: poses into pose tables and FP64 force rows into
engine loads against the coupling oracle, the free-body integrator against
closed forms, and a two-way coupled sphere released in a channel flow."""
import math

import numpy as np
import pytest
import torch

from oracle import coupling as O
from openhomelbm_amd import _native as N
from openhomelbm_amd import coupling as CP
from openhomelbm_amd.lattice import Lattice

pytestmark = pytest.mark.gpu
F = np.float32


def _quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


@pytest.mark.parametrize("lattice", [9, 27])
def test_poses_into_tables(lattice):
    rng = np.random.default_rng(71)
    nworld, nbody, nsolid = 3, 2, 3
    lats = [Lattice(lattice, 24, 20, 1 if lattice == 9 else 16, nsolid=nsolid) for _ in range(nworld)]
    bm = CP.BodyMap([2, 0], engine_origin=[[0.5, -0.25, 1.0], [0, 0, 0]], lattice_origin=[[10, 8, 6], [3, 4, 5]],
                    scale=[16.0, 8.0])
    init = torch.zeros((nworld, nsolid), dtype=torch.int32, device="cuda")
    width = 8 if lattice == 9 else 24
    start = [lat.solids.cpu().numpy()[:nsolid].reshape(nsolid, width).copy() for lat in lats]
    poses = []
    for call in range(2):
        pos = (rng.random((nworld, nbody, 3)) - 0.5).astype(F)
        quat = _quats(rng, nworld * nbody).reshape(nworld, nbody, 4).astype(F)
        CP.poses_to_lattices(lats, bm, torch.from_numpy(pos).cuda(), torch.from_numpy(quat).cuda(), init)
        torch.cuda.synchronize()
        poses.append((pos, quat))
        eo, lo, sc = bm.engine_origin.cpu().numpy(), bm.lattice_origin.cpu().numpy(), bm.scale.cpu().numpy()
        for w in range(nworld):
            T = lats[w].solids.cpu().numpy()[:nsolid].reshape(nsolid, width)
            for b, sid in enumerate([2, 0]):
                want = [O.lattice_position(pos[w, b, k], eo[b, k], lo[b, k], sc[b]) for k in range(3)]
                half = width // 2
                if lattice == 9:
                    ang = O.planar_angle(quat[w, b])
                    assert T[sid, 0] == want[0] and T[sid, 1] == want[1]
                    np.testing.assert_allclose(T[sid, 2:4], [math.cos(ang), math.sin(ang)], atol=3e-7)
                else:
                    assert np.array_equal(T[sid, 0:3], np.array(want, F))
                    np.testing.assert_allclose(T[sid, 3:12], O.quat_matrix(quat[w, b]), atol=1e-7)
                prev = T[sid, :half] if call == 0 else prev_tables[w][sid, :half]
                assert np.array_equal(T[sid, half:], prev)  # last = current at first, then the previous pose
            assert np.array_equal(T[1], start[w][1])  # solid 1 has no body: untouched
        prev_tables = [lats[w].solids.cpu().numpy()[:nsolid].reshape(nsolid, width).copy() for w in range(nworld)]
    assert init.cpu().numpy()[:, [0, 2]].all() and not init.cpu().numpy()[:, 1].any()


@pytest.mark.parametrize("lattice", [9, 27])
@pytest.mark.parametrize("mode", [N.HL_COUPLE_SCALED, N.HL_COUPLE_PHYSICAL])
def test_force_rows_into_engine_loads(lattice, mode):
    rng = np.random.default_rng(72)
    nworld, nsolid = 4, 3
    rows = rng.standard_normal((nworld, nsolid, 6))
    bm = CP.BodyMap([1, 2], scale=[2.5, 4.0], body_ids=[5, 0])
    xfrc = torch.full((nworld, 7, 6), -7.0, dtype=torch.float32, device="cuda")
    f, t = CP.forces_to_engine(lattice, torch.from_numpy(rows).cuda(), nworld, nsolid, bm, mode, 0.75, 1.5,
                               xfrc=xfrc)
    f, t, x = f.cpu().numpy(), t.cpu().numpy(), xfrc.cpu().numpy()
    for w in range(nworld):
        for b, (sid, s, body) in enumerate([(1, 2.5, 5), (2, 4.0, 0)]):
            wf, wt = O.engine_loads(rows[w, sid], lattice, mode, s, 0.75, 1.5)
            assert np.array_equal(f[w, b], wf) and np.array_equal(t[w, b], wt)
            assert np.array_equal(x[w, body], np.concatenate([wf, wt]))
        assert np.all(x[w, [1, 2, 3, 4, 6]] == -7.0)  # other bodies untouched


def test_free_body_closed_forms():
    n, dt = 40, 0.5
    m, I = 3.0, 2.0
    fb = CP.FreeBodies(2, [[[1.0, 2.0, 3.0]], [[0.0, 0.0, 0.0]]], [m], [[I, I, I]])
    # world 1 spins about z at 0.1 rad per unit time with no torque
    s = fb.state.cpu().numpy()
    s[1, 0, 12] = 0.1
    fb.state.copy_(torch.from_numpy(s))
    force = torch.tensor([[[0.3, -0.6, 0.15]], [[0.0, 0.0, 0.0]]], dtype=torch.float32, device="cuda")
    torque = torch.tensor([[[0.0, 0.0, 0.04]], [[0.0, 0.0, 0.0]]], dtype=torch.float32, device="cuda")
    for _ in range(n):
        fb.step(force, torque, dt, gravity=(0.0, 0.0, -0.01))
    s = fb.state.cpu().numpy().astype(np.float64)
    a = np.array([0.3, -0.6, 0.15]) / m + np.array([0, 0, -0.01])
    # semi-implicit Euler: v_n = n dt a, p_n = p_0 + dt^2 a n (n + 1) / 2
    np.testing.assert_allclose(s[0, 0, 7:10], n * dt * a, rtol=1e-5)
    np.testing.assert_allclose(s[0, 0, 0:3], [1, 2, 3] + dt * dt * a * n * (n + 1) / 2, rtol=1e-5)
    np.testing.assert_allclose(s[0, 0, 12], n * dt * 0.04 / I, rtol=1e-5)  # isotropic: omega_n = n dt tau / I
    # torque-free spin about z: the half angle advances by atan(omega dt / 2) per step
    q = s[1, 0, 3:7]
    np.testing.assert_allclose(np.linalg.norm(q), 1.0, atol=1e-6)
    half = n * math.atan(0.1 * dt / 2)
    np.testing.assert_allclose(q, [math.cos(half), 0, 0, math.sin(half)], atol=2e-6)


def test_body_capsules_follow_pose():
    fb = CP.FreeBodies(1, [[[2.0, 1.0, 0.5]]], [1.0], [[1, 1, 1]],
                       quat=[[[math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)]]])
    bm = CP.BodyMap([0], engine_origin=[[1.0, 1.0, 0.0]], lattice_origin=[[32, 16, 8]], scale=[10.0])
    shapes = torch.tensor([[-0.5, 0, 0, 0.5, 0, 0, 0.2]], dtype=torch.float32, device="cuda")
    caps = fb.capsules(bm, shapes).cpu().numpy()[0, 0]
    # body x axis turned onto world y: ends at (2, 1 -/+ 0.5, 0.5) -> lattice (42, 16 -/+ 5, 13)
    np.testing.assert_allclose(caps[:6], [42, 11, 13, 42, 21, 13], atol=1e-5)
    assert caps[6] == F(2.0) and caps[7] == 0


def _channel(nsolid=1):
    lat = Lattice(27, 64, 32, 32, nsolid=nsolid)
    lat.nu = 0.02
    lat.bc_type = [0, 1, 1, 1, 1, 1]
    lat.bc_value = [[0.04, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
    lat.init_uniform((0.04, 0.0, 0.0))
    return lat


def _released_sphere(steps):
    r = 4.0
    mass = 2.0 * 4.0 / 3.0 * math.pi * r ** 3   # twice the displaced fluid
    lat = _channel()
    fb = CP.FreeBodies(1, [[[20.0, 16.0, 16.0]]], [mass], [[0.4 * mass * r * r] * 3])
    bm = CP.BodyMap([0])
    world = CP.CoupledWorlds([lat], fb, bm, [[0, 0, 0, 0, 0, 0, r]], box=(10, 44, 8, 24, 8, 24))
    vx = []
    for _ in range(steps // 50):
        world.step(50)
        vx.append(float(fb.state[0, 0, 7]))
    return lat, fb, vx


def test_released_sphere_is_carried_by_the_flow():
    lat, fb, vx = _released_sphere(300)
    s = fb.state.cpu().numpy()[0, 0]
    assert np.isfinite(lat.get_state(np.float32)).all()
    assert all(b > a for a, b in zip(vx, vx[1:])), vx      # accelerates downstream
    assert 0.0 < vx[-1] < 0.04, vx                          # but slower than the stream
    assert abs(s[8]) < 0.1 * vx[-1] and abs(s[9]) < 0.1 * vx[-1]
    assert s[0] > 20.0
    # deterministic: the same run again is bitwise equal
    lat2, fb2, _ = _released_sphere(300)
    assert torch.equal(fb.state, fb2.state)
    assert np.array_equal(lat.get_state(np.float32), lat2.get_state(np.float32))
