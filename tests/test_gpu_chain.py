"""Planar chains on the GPU (hl_chain_step) against the chain oracle, and the
coupled eel loop (ChainWorlds): the reference's _simulation_step
(envs/lbm/lbm_fluid_env.py:441-533) with the chain in MuJoCo's place.

Bars: FP64 joint states within 1e-10 relative of the oracle after tens of
steps (the kernel and the oracle differ in summation order, the solve and
the libm trigonometry, so not bitwise); FP32 poses within 2 ulp-scale of the
oracle's FP64 poses; split calls bitwise equal to one call; the coupled loop
deterministic, its chain states reproduced by the oracle fed the loop's own
loads, and fluid drag slowing a coasting eel.
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import chain as oc
from openhomelbm_amd.chain import Chain, ChainWorlds, chain_from_mjcf
from openhomelbm_amd.lattice import Lattice
from tests.test_chain_oracle import to_oracle, tree

pytestmark = pytest.mark.gpu
EEL = os.path.join(os.path.dirname(__file__), "golden", "eel_2d.xml")


def _oracle_run(om, q, v, ctrl, xfrc, n, factor):
    applied = None
    for _ in range(n):
        q, v, applied = oc.step(om, q, v, ctrl, xfrc, applied_factor=factor)
    return q, v, applied


def _check(ch, om, q0, v0, ctrl, xfrc, n, factor=1.0):
    for w in range(ch.nworld):
        q, v, applied = _oracle_run(om, q0[w], v0[w], ctrl[w].astype(np.float64) if om.act_dof.size else [],
                                    xfrc[w].astype(np.float64), n, factor)
        gq, gv = ch.qpos[w].cpu().numpy(), ch.qvel[w].cpu().numpy()
        np.testing.assert_allclose(gq, q, rtol=1e-10, atol=1e-12, err_msg=f"world {w} qpos")
        np.testing.assert_allclose(gv, v, rtol=1e-10, atol=1e-11, err_msg=f"world {w} qvel")
        np.testing.assert_allclose(ch.qfrc_applied[w].cpu().numpy(), applied, rtol=1e-10, atol=1e-12)
        xpos, xipos, xquat = oc.poses(om, q)
        np.testing.assert_allclose(ch.xpos[w].cpu().numpy(), xpos, rtol=3e-7, atol=3e-7)
        np.testing.assert_allclose(ch.xipos[w].cpu().numpy(), xipos, rtol=3e-7, atol=3e-7)
        np.testing.assert_allclose(np.abs(ch.xquat[w].cpu().numpy() @ xquat.T).diagonal(), 1.0, atol=3e-7)


def test_eel_steps_match_oracle():
    rng = np.random.default_rng(21)
    m = chain_from_mjcf(EEL)
    om = to_oracle(m)
    nworld = 3
    ch = Chain(m, nworld)
    q0 = m.qpos0 + rng.uniform(-0.3, 0.3, (nworld, m.nv))
    v0 = rng.uniform(-0.5, 0.5, (nworld, m.nv))
    ch.qpos.copy_(torch.from_numpy(q0))
    ch.qvel.copy_(torch.from_numpy(v0))
    ctrl = rng.uniform(-1.5, 1.5, (nworld, m.nact)).astype(np.float32)  # some beyond the ctrlrange
    xfrc = (rng.normal(size=(nworld, m.nbody, 6)) * [0.05, 0.05, 1.0, 1.0, 1.0, 0.005]).astype(np.float32)
    ch.step(25, ctrl=torch.from_numpy(ctrl).cuda(), xfrc=torch.from_numpy(xfrc).cuda())
    torch.cuda.synchronize()
    _check(ch, om, q0, v0, ctrl, xfrc, 25)


def test_tree_euler_gravity_stiffness_and_no_applied():
    rng = np.random.default_rng(4)
    om = tree()
    om.integrator = "Euler"
    om.gravity = np.array([0.3, -9.81])
    om.damping = rng.uniform(0.0, 0.2, om.nv)
    om.stiffness = rng.uniform(0.0, 1.0, om.nv)
    om.armature = rng.uniform(0.0, 0.05, om.nv)
    om.springref = rng.uniform(-0.2, 0.2, om.nv)
    om.act_dof, om.gear = np.array([3, 5]), np.array([0.7, -0.3])
    om.ctrl_lo, om.ctrl_hi, om.ctrl_limited = np.array([-1.0, -1.0]), np.array([1.0, 1.0]), np.array([False, True])
    from openhomelbm_amd.chain import ChainModel
    pm = ChainModel(**{f: getattr(om, f) for f in (
        "parent", "free", "pos", "angle", "com", "mass", "inertia", "posz", "comz", "dof_body", "dof_type",
        "dof_axis", "dof_anchor", "dof_sign", "damping", "stiffness", "springref", "armature", "qpos0", "act_dof",
        "gear", "ctrl_lo", "ctrl_hi", "ctrl_limited", "gravity", "dt", "integrator")}, applied_factor=0.0)
    nworld = 2
    ch = Chain(pm, nworld)
    q0, v0 = rng.uniform(-0.5, 0.5, (nworld, om.nv)), rng.uniform(-0.5, 0.5, (nworld, om.nv))
    ch.qpos.copy_(torch.from_numpy(q0))
    ch.qvel.copy_(torch.from_numpy(v0))
    ctrl = rng.uniform(-2, 2, (nworld, 2)).astype(np.float32)
    xfrc = (rng.normal(size=(nworld, om.nbody, 6)) * 0.1).astype(np.float32)
    ch.step(40, ctrl=torch.from_numpy(ctrl).cuda(), xfrc=torch.from_numpy(xfrc).cuda())
    torch.cuda.synchronize()
    _check(ch, om, q0, v0, ctrl, xfrc, 40, factor=0.0)


def test_split_calls_equal_one_call():
    m = chain_from_mjcf(EEL)
    a, b = Chain(m, 2), Chain(m, 2)
    ctrl = torch.tensor([[0.3, -0.2, 0.5, 0.1, -0.9], [1.0, 0.0, -1.0, 0.4, 0.2]], dtype=torch.float32,
                        device="cuda")
    xfrc = torch.zeros((2, m.nbody, 6), dtype=torch.float32, device="cuda")
    xfrc[:, :, 0] = 0.02
    a.step(12, ctrl=ctrl, xfrc=xfrc)
    for _ in range(3):
        b.step(4, ctrl=ctrl, xfrc=xfrc)
    torch.cuda.synchronize()
    assert torch.equal(a.qpos, b.qpos) and torch.equal(a.qvel, b.qvel) and torch.equal(a.xipos, b.xipos)


# ----------------------------------------------------------- coupled eel

SCALE = 110.0  # lattice cells per engine metre


def _eel_world(v_init=None):
    m = chain_from_mjcf(EEL)
    ch = Chain(m, 1)
    ch.qpos[0, 0:3] = torch.tensor([0.0, 0.0, math.pi / 2], dtype=torch.float64)  # eel along +x from its head
    if v_init is not None:
        ch.qvel[0, 0:2] = torch.tensor(v_init, dtype=torch.float64)
    ch.step(0)
    lat = Lattice(9, 176, 80, nsolid=m.nbody)
    lat.nu = 0.02
    lat.bc_type = [0, 0, 0, 0]
    lat.bc_value = [[0.0, 0.0, 0.0]] * 4
    lat.init_uniform((0.0, 0.0, 0.0))
    bodies = list(range(m.nbody))
    world = ChainWorlds([lat], ch, bodies, bodies, box=(10, 166, 10, 70),
                        engine_origin=[[0.0, 0.0, 0.05]] * m.nbody, lattice_origin=[[50.0, 40.0, 0.0]] * m.nbody,
                        scale=[SCALE] * m.nbody)
    return m, ch, lat, world


def _ctrl(k):
    ph = 2 * math.pi * k / 60.0
    return torch.tensor([[0.15 * math.sin(ph - 0.8 * j) for j in range(5)]], dtype=torch.float32, device="cuda")


def test_coupled_eel_is_deterministic_and_replays_in_the_oracle():
    m, ch, lat, world = _eel_world()
    om = to_oracle(m)
    q0, v0 = ch.qpos[0].cpu().numpy().copy(), ch.qvel[0].cpu().numpy().copy()
    loads, ctrls = [], []
    for k in range(30):
        c = _ctrl(k)
        world.step(1, ctrl=c)
        loads.append(ch.xfrc[0].cpu().numpy().astype(np.float64))
        ctrls.append(c[0].cpu().numpy().astype(np.float64))
    torch.cuda.synchronize()
    assert np.isfinite(lat.get_state(np.float32)).all()
    assert np.abs(np.array(loads)).max() > 0.0  # the fluid pushes back
    q, v = q0, v0
    for L, c in zip(loads, ctrls):
        q, v, _ = oc.step(om, q, v, c, L)
    np.testing.assert_allclose(ch.qpos[0].cpu().numpy(), q, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(ch.qvel[0].cpu().numpy(), v, rtol=1e-9, atol=1e-11)
    # the same run again: bitwise
    _, ch2, lat2, world2 = _eel_world()
    for k in range(30):
        world2.step(1, ctrl=_ctrl(k))
    torch.cuda.synchronize()
    assert torch.equal(ch.qpos, ch2.qpos) and torch.equal(lat.fields(), lat2.fields())


def test_fluid_drag_slows_a_coasting_eel():
    v = (-0.04, 0.0)  # head first (the eel lies along +x from its head)
    m, ch, lat, world = _eel_world(v_init=v)
    zero = torch.zeros((1, m.nact), dtype=torch.float32, device="cuda")
    world.step(40, ctrl=zero)
    free = Chain(m, 1)
    free.qpos.copy_(ch.qpos.new_tensor([[0.0, 0.0, math.pi / 2] + [0.0] * 5]))
    free.qvel[0, 0:2] = torch.tensor(v, dtype=torch.float64)
    free.step(40, ctrl=zero)
    torch.cuda.synchronize()
    vc, vf = ch.qvel[0, 0].item(), free.qvel[0, 0].item()
    assert vf < 0.0 and vc < 0.0
    # the fluid takes away several times what joint damping alone does
    assert abs(v[0]) - abs(vc) > 3.0 * (abs(v[0]) - abs(vf)) > 0.0, (vc, vf)
