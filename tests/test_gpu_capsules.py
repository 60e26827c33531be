"""GPU parity of the capsule cut-link step (hl_step_capsules) against the
oracle's hlo_step_capsules: the ray query of get_cutcell_multi_3d
(envs/lbm3d/lbm_func_3d.py:254-316) with closed-form capsule surfaces and the
wall rule at the crossing (lbm_func_3d.py:424-511).

Bars, all written in the tests:
  * gather kernel + near-body pass (FP32 and FP64): the whole field equals the
    oracle bit for bit, step after step;
  * tiled kernel (FP32): every cell with a cut link (the near-body pass owns
    it) equals the oracle bit for bit; the field within relative L2 1e-6 per
    step (the tiled kernel's arithmetic, as for the mask rule);
  * forces: FP64 sums of bitwise-equal FP32 link forces in another order,
    rtol 1e-12.
The full-size sphere (256x128x128) runs one tiled step against one OpenMP
oracle step, with the mask rule and with capsule crossings.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from openhomelbm_amd import _native as N
from openhomelbm_amd import inputs
from openhomelbm_amd.lattice import Lattice, capsule_box
from openhomelbm_amd.lbm3d import LBM_Solver3D
from tests.helpers import random_state, tensor_rel_l2

pytestmark = pytest.mark.gpu

SHAPE = (34, 40, 72)  # nz, ny, nx


def _scene(seed):
    """A static sphere (solid 0), a moving, turning capsule (solid 1) and a
    second capsule of solid 1 crossing the first (its surfaces count on their
    own), plus a body-force field."""
    nz, ny, nx = SHAPE
    rng = np.random.default_rng(seed)
    caps = np.array([[20.3, 18.6, 15.2, 20.3, 18.6, 15.2, 4.4, 0],
                     [44.2, 21.1, 17.4, 55.6, 26.3, 20.8, 2.7, 1],
                     [49.0, 17.5, 22.0, 51.0, 29.0, 15.5, 2.2, 1]], np.float32)
    ang = 0.03
    c, s = np.cos(ang), np.sin(ang)
    poses = np.zeros((2, 24), np.float32)
    poses[:, 3:12] = np.eye(3).ravel()
    poses[:, 15:24] = np.eye(3).ravel()
    poses[0, 0:3] = poses[0, 12:15] = (20.3, 18.6, 15.2)
    poses[1, 0:3] = (49.9, 23.7, 19.1)
    poses[1, 3:12] = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]).ravel()
    poses[1, 12:15] = (49.87, 23.72, 19.1)
    env = orc.Env(27, nx, ny, nz, nu=0.02, bc_type=(0, 2, 1, 1, 2, 1),
                  bc_value=((0.05, 0.01, 0.005),) + ((0.0, 0.0, 0.0),) * 5)
    env.solids = poses
    env.force = 2e-4 * rng.standard_normal((3,) + SHAPE).astype(np.float32)
    return env, caps, poses


def _lattice(env, caps, poses, dtype="f32"):
    lat = Lattice(27, env.nx, env.ny, env.nz, dtype=dtype, nsolid=poses.shape[0])
    lat.nu = env.nu
    lat.bc_type = list(env.bc_type)
    lat.bc_value = [list(v) for v in env.bc_value]
    lat.set_poses(poses)
    lat.set_force(env.force)
    caps_d = torch.from_numpy(caps).to(lat.device)
    box = capsule_box(caps, (env.nx, env.ny, env.nz))
    return lat, caps_d, box


def _cut_cells(env, caps):
    """Cells whose interior projection has a capsule cut link (the cells the
    near-body pass owns), from the oracle's crossing query."""
    near = np.zeros(env.shape, bool)
    box = capsule_box(caps, (env.nx, env.ny, env.nz))
    for z in range(max(box[4] - 2, 1), min(box[5] + 2, env.nz - 1)):
        for y in range(max(box[2] - 2, 1), min(box[3] + 2, env.ny - 1)):
            for x in range(max(box[0] - 2, 1), min(box[1] + 2, env.nx - 1)):
                near[z, y, x] = any(orc.capsule_cut(env, caps, (x, y, z), i)[0] >= 0 for i in range(1, 27))
    return near


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_gather_capsule_steps_equal_the_oracle(dtype):
    env, caps, poses = _scene(3)
    np_t = np.float32 if dtype == "f32" else np.float64
    m = random_state(27, SHAPE, seed=51).astype(np_t)
    lat, caps_d, box = _lattice(env, caps, poses, dtype)
    lat.set_state(m)
    n = 3
    h = torch.zeros((n, 2, 6), dtype=torch.float64, device=lat.device)
    lat.step_capsules(n, caps_d, box, kernel=N.HL_KERNEL_GATHER, force_history=h)
    hist = h.cpu().numpy()
    ref = m
    for k in range(n):
        ref, f = orc.step_capsules(env, ref, caps, dtype)
        assert np.allclose(hist[k], f, rtol=1e-12, atol=1e-14), (k, hist[k], f)
    got = lat.get_state(np_t)
    print(f"{dtype}: max |diff| {np.abs(got - ref).max()}")
    assert np.array_equal(got, ref)


def test_tiled_capsule_steps_match_the_oracle():
    env, caps, poses = _scene(4)
    m = random_state(27, SHAPE, seed=52).astype(np.float32)
    lat, caps_d, box = _lattice(env, caps, poses)
    near = _cut_cells(env, caps)
    assert near.sum() > 2000
    ref = m
    for k in range(3):
        lat.set_state(ref)
        h = torch.zeros((1, 2, 6), dtype=torch.float64, device=lat.device)
        lat.step_capsules(1, caps_d, box, kernel=N.HL_KERNEL_TILED, force_history=h)
        got = lat.get_state(np.float32)
        ref_next, f = orc.step_capsules(env, ref, caps, "f32")
        assert np.allclose(h.cpu().numpy()[0], f, rtol=1e-12, atol=1e-14), (h, f)
        assert np.array_equal(got[:, near], ref_next[:, near]), k
        e = tensor_rel_l2(27, got, ref_next)
        print(f"step {k}: {e}")
        assert e["all"] < 1e-6, e
        assert not np.array_equal(got, ref_next)  # cells away from the bodies took the tiled path
        ref = ref_next


def test_tiled_multi_step_call_equals_single_steps():
    """One call of 4 steps (the list built once, forces of each step summed by
    the next launch) equals 4 calls of one step, states and forces bitwise."""
    env, caps, poses = _scene(5)
    m = random_state(27, SHAPE, seed=53).astype(np.float32)
    outs = []
    for per_call in (False, True):
        lat, caps_d, box = _lattice(env, caps, poses)
        lat.set_state(m)
        h = torch.zeros((4, 2, 6), dtype=torch.float64, device=lat.device)
        if per_call:
            for k in range(4):
                lat.step_capsules(1, caps_d, box, force_history=h[k:k + 1])
        else:
            lat.step_capsules(4, caps_d, box, force_history=h)
        outs.append((lat.get_state(np.float32), h.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


def test_capsule_step_rejects_bad_arguments():
    env, caps, poses = _scene(6)
    lat, caps_d, box = _lattice(env, caps, poses)
    with pytest.raises(N.HLError):
        lat.step_capsules(1, caps_d, (0, env.nx + 1, 0, env.ny, 0, env.nz))
    lat2 = Lattice(9, 32, 32, nsolid=1)
    with pytest.raises(ValueError):
        lat2.step_capsules(1, caps_d, box)


def test_solver_cut_cell_mode_takes_capsule_crossings():
    """LBM_Solver3D(cut_cell=True) with a sphere body steps through the
    capsule cut links: its forces equal those of Lattice.step_capsules on the
    same state, and differ from the mask rule's."""
    nx, ny, nz = 64, 40, 40
    out = {}
    for cut in (True, False):
        sol = LBM_Solver3D(nx, ny, nz, solid_num=1, cut_cell=cut)
        f = sol.flows[0]
        f.vis_shear = 0.02
        f.bc_type = np.array([0, 1, 1, 1, 1, 1])
        f.bc_value = np.array([[0.04, 0, 0]] + [[0, 0, 0]] * 5)
        sol.create_solid_sphere(0, 6.3, lbm_position=(20.4, 19.7, 20.2), lbm_scale=1.0 / nx)
        for _ in range(3):
            sol.step()
        out[cut] = (f.solid_force.numpy().copy(), f.solid_torque.numpy().copy(), f.lattice.get_state(np.float32),
                    f.capsules.copy())
    F, T, state, caps = out[True]
    assert not np.allclose(F, out[False][0])
    # the same three steps through the lattice entry point
    lat = Lattice(27, nx, ny, nz, nsolid=1)
    lat.nu = 0.02
    lat.bc_type = [0, 1, 1, 1, 1, 1]
    lat.bc_value = [[0.04, 0, 0]] + [[0, 0, 0]] * 5
    pose = np.zeros((1, 24), np.float32)
    pose[0, 0:3] = pose[0, 12:15] = (20.4, 19.7, 20.2)
    pose[0, 3:12] = pose[0, 15:24] = np.eye(3).ravel()
    lat.set_poses(pose)
    lat.init_uniform((0.0, 0.0, 0.0))
    caps_d = torch.from_numpy(caps).to(lat.device)
    h = torch.zeros((3, 1, 6), dtype=torch.float64, device=lat.device)
    box = capsule_box(caps, (nx, ny, nz))
    for k in range(3):
        lat.step_capsules(1, caps_d, box, force_history=h[k:k + 1])
    assert np.array_equal(lat.get_state(np.float32), state)
    hh = h.cpu().numpy()
    # solid_force / solid_torque hold FP32 vectors, as the reference's wp.vec3 arrays do
    assert np.array_equal(hh[2, 0, 0:3].astype(np.float32), F[0])
    assert np.array_equal(hh[2, 0, 3:6].astype(np.float32), T[0])


def _full_env(w):
    env = orc.Env(27, w.nx, w.ny, w.nz, nu=w.nu, bc_type=w.bc_type, bc_value=w.bc_value)
    env.solids = w.poses
    return env


@pytest.mark.parametrize("boundary", ["mask", "capsules"])
def test_full_size_sphere_tiled_vs_openmp_oracle(boundary):
    """The 256x128x128 sphere at Re 200: one tiled FP32 step against one
    OpenMP FP32 oracle step from the same state, with the mask rule and with
    capsule crossings: cells with a cut link bit for bit, the field within
    relative L2 1e-6."""
    w = inputs.d3q27_sphere()
    lat = inputs.build_lattice(w)
    m0 = lat.get_state(np.float32)
    env = _full_env(w)
    if boundary == "mask":
        env.mask = lat.mask.view(w.nz, w.ny, w.nx).cpu().numpy()
        lat.step(1, kernel=N.HL_KERNEL_TILED)
        ref = orc.step(env, m0, "f32", omp=True)
        occ = env.mask != 0
        near = np.zeros_like(occ)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    near |= np.roll(occ, (dz, dy, dx), axis=(0, 1, 2)) != occ
    else:
        caps_d = torch.from_numpy(w.capsules).to(lat.device)
        lat.step_capsules(1, caps_d, capsule_box(w.capsules, (w.nx, w.ny, w.nz)), kernel=N.HL_KERNEL_TILED)
        ref, _ = orc.step_capsules(env, m0, w.capsules, omp=True)
        near = _cut_cells(env, w.capsules)
    got = lat.get_state(np.float32)
    e = tensor_rel_l2(27, got, ref)
    print(f"{boundary}: {int(near.sum())} near-body cells, {e}")
    assert near.sum() > 5000
    assert np.array_equal(got[:, near], ref[:, near])
    assert e["all"] < 1e-6, e
