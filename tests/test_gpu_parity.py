"""GPU parity: the HIP kernels through the C ABI against the CPU oracle on the
same seeded inputs.

* The gather kernel evaluates the reference's operation order without
  contraction, so its FP32 build must equal the FP32 oracle bit for bit and
  its FP64 build the FP64 oracle.
* The voxelizer is bit-exact; forces agree to FP64 summation-order rounding
  (relative 1e-12); the force reduction is bitwise repeatable.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from openhomelbm_amd import _native as N
from openhomelbm_amd.lattice import Lattice
from tests.helpers import random_state, tensor_rel_l2

pytestmark = pytest.mark.gpu


def _scene(lattice, shape, seed, moving=True, with_force=True):
    """Oracle env + GPU lattice with the same bodies, BCs, force field."""
    rng = np.random.default_rng(seed)
    d = 2 if lattice == 9 else 3
    if d == 2:
        ny, nx = shape
        nz = 1
        caps = np.array([[nx * 0.3, ny * 0.5, 0, nx * 0.3, ny * 0.5, 0, 4.2, 0],
                         [nx * 0.55, ny * 0.3, 0, nx * 0.75, ny * 0.62, 0, 2.6, 1]], np.float32)
        ang = 0.05 if moving else 0.0
        poses = np.array([[nx * 0.3, ny * 0.5, 1, 0, nx * 0.3, ny * 0.5, 1, 0],
                          [nx * 0.65, ny * 0.46, np.cos(ang), np.sin(ang),
                           nx * 0.65 - (0.02 if moving else 0), ny * 0.46 + (0.01 if moving else 0), 1, 0]],
                         np.float32)
        bc_type = (0, 2, 1, 1)
    else:
        nz, ny, nx = shape
        caps = np.array([[nx * 0.3, ny * 0.5, nz * 0.5, nx * 0.3, ny * 0.5, nz * 0.5, 3.4, 0],
                         [nx * 0.55, ny * 0.3, nz * 0.4, nx * 0.72, ny * 0.62, nz * 0.6, 2.2, 1]], np.float32)
        ang = 0.04 if moving else 0.0
        c, s = np.cos(ang), np.sin(ang)
        poses = np.zeros((2, 24), np.float32)
        poses[0, 0:3] = [nx * 0.3, ny * 0.5, nz * 0.5]
        poses[0, 12:15] = poses[0, 0:3]
        poses[:, 3:12] = np.eye(3).ravel()
        poses[:, 15:24] = np.eye(3).ravel()
        poses[1, 0:3] = [nx * 0.63, ny * 0.46, nz * 0.5]
        poses[1, 3:12] = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]).ravel()
        poses[1, 12:15] = poses[1, 0:3] - (np.array([0.02, -0.01, 0.005]) if moving else 0)
        bc_type = (0, 2, 1, 1, 2, 1)
    bc_value = ((0.05, 0.01, 0.005),) + ((0.0, 0.0, 0.0),) * 5
    nu = 0.02
    env = orc.Env(lattice, nx, ny, nz, nu=nu, bc_type=bc_type, bc_value=bc_value)
    mask = orc.voxelize(env, caps)
    env.mask = mask
    env.solids = poses
    force = 2e-4 * rng.standard_normal((d,) + tuple(shape)).astype(np.float32) if with_force else None
    env.force = force
    return env, caps, poses, force


def _gpu_lattice(env, caps, poses, force, dtype="f32"):
    lat = Lattice(env.lattice, env.nx, env.ny, env.nz, dtype=dtype, nsolid=poses.shape[0])
    lat.nu = env.nu
    lat.bc_type = list(env.bc_type) + [1] * (6 - len(env.bc_type))
    lat.bc_value = [list(v) for v in env.bc_value]
    lat.set_poses(poses)
    lat.voxelize(caps)
    lat.set_force(force)
    return lat


CASES = [(9, (40, 56)), (27, (18, 22, 26))]


@pytest.mark.parametrize("lattice,shape", CASES)
def test_voxelizer_bit_exact(lattice, shape):
    env, caps, poses, force = _scene(lattice, shape, 1)
    lat = _gpu_lattice(env, caps, poses, force)
    torch.cuda.synchronize()
    got = lat.mask.cpu().numpy().reshape(env.shape)
    assert np.array_equal(got, env.mask)
    assert got.any()


def test_voxelizer_sphere_closed_form_full_size():
    """The 256x128x128 sphere of diameter 24 at (64, 64, 65): integer closed form."""
    lat = Lattice(27, 256, 128, 128, nsolid=1)
    lat.voxelize(np.array([[64, 64, 65, 64, 64, 65, 12, 0]], np.float32))
    m = lat.mask.view(128, 128, 256).cpu().numpy()
    z, y, x = np.ogrid[0:128, 0:128, 0:256]
    expect = ((x - 64) ** 2 + (y - 64) ** 2 + (z - 65) ** 2 <= 144)
    assert np.array_equal(m != 0, expect)
    assert lat.mask_box[0] <= 52 and lat.mask_box[1] >= 77


@pytest.mark.parametrize("lattice,shape", CASES)
def test_init_uniform_bit_exact(lattice, shape):
    env = orc.Env(lattice, shape[-1], shape[-2], shape[0] if lattice == 27 else 1)
    u = (0.05, -0.01, 0.02)
    lat = Lattice(lattice, env.nx, env.ny, env.nz)
    lat.init_uniform(u)
    got = lat.get_state(np.float32)
    assert np.array_equal(got, orc.init_uniform(env, u, "f32"))


@pytest.mark.parametrize("lattice,shape", CASES)
@pytest.mark.parametrize("with_body", [False, True])
@pytest.mark.parametrize("nsteps", [1, 4])
def test_gather_f32_bit_exact(lattice, shape, with_body, nsteps):
    env, caps, poses, force = _scene(lattice, shape, 2)
    if not with_body:
        env.mask = None
        env.solids = None
    m0 = random_state(lattice, shape, seed=31).astype(np.float32)
    lat = _gpu_lattice(env, caps, poses, force)
    if not with_body:
        lat.set_mask(None)
    lat.set_state(m0)
    lat.step(nsteps, kernel=N.HL_KERNEL_GATHER)
    got = lat.get_state(np.float32)
    ref = m0
    for _ in range(nsteps):
        ref = orc.step(env, ref, "f32")
    if not np.array_equal(got, ref):
        err = tensor_rel_l2(lattice, got, ref)
        nbad = int((got != ref).sum())
        pytest.fail(f"not bit-exact: {nbad} values differ, rel-L2 {err}")


@pytest.mark.parametrize("lattice,shape", CASES)
def test_gather_f64_matches_fp64_oracle(lattice, shape):
    env, caps, poses, force = _scene(lattice, shape, 3)
    m0 = random_state(lattice, shape, seed=32)
    lat = _gpu_lattice(env, caps, poses, force, dtype="f64")
    lat.set_state(m0)
    lat.step(3, kernel=N.HL_KERNEL_GATHER)
    got = lat.get_state()
    ref = m0
    for _ in range(3):
        ref = orc.step(env, ref, "f64")
    err = tensor_rel_l2(lattice, got, ref)
    assert max(err.values()) < 1e-14, err


@pytest.mark.parametrize("lattice,shape", CASES)
def test_forces_match_oracle_and_repeat_bitwise(lattice, shape):
    env, caps, poses, force = _scene(lattice, shape, 4)
    m0 = random_state(lattice, shape, seed=33).astype(np.float32)
    lat = _gpu_lattice(env, caps, poses, force)
    lat.set_state(m0)
    f1 = lat.forces().cpu().numpy().copy()
    f2 = lat.forces().cpu().numpy().copy()
    assert np.array_equal(f1, f2)  # fixed-order FP64 reduction
    ref = orc.compute_forces(env, m0, "f32")
    assert np.allclose(f1, ref, rtol=1e-12, atol=1e-14), (f1, ref)
    assert np.abs(ref[:, :2]).max() > 1e-5


@pytest.mark.parametrize("lattice,shape", CASES)
def test_force_history_during_steps(lattice, shape):
    env, caps, poses, force = _scene(lattice, shape, 5)
    m0 = random_state(lattice, shape, seed=34).astype(np.float32)
    lat = _gpu_lattice(env, caps, poses, force)
    lat.set_state(m0)
    hist = torch.zeros((3, 2, 6), dtype=torch.float64, device=lat.device)
    lat.step(3, kernel=N.HL_KERNEL_GATHER, force_history=hist)
    h = hist.cpu().numpy()
    cur = m0
    for k in range(3):
        assert np.allclose(h[k], orc.compute_forces(env, cur, "f32"), rtol=1e-12, atol=1e-14)
        cur = orc.step(env, cur, "f32")


@pytest.mark.parametrize("lattice,shape", CASES)
def test_mass_fp64(lattice, shape):
    env = orc.Env(lattice, shape[-1], shape[-2], shape[0] if lattice == 27 else 1)
    m0 = random_state(lattice, shape, seed=35).astype(np.float32)
    lat = Lattice(lattice, env.nx, env.ny, env.nz)
    lat.set_state(m0)
    assert abs(lat.mass() - m0[0].astype(np.float64).sum()) < 1e-9


def test_invalid_box_raises():
    lat = Lattice(9, 16, 12, nsolid=1)
    with pytest.raises(N.HLError):
        lat.voxelize(np.zeros((1, 8), np.float32), box=(0, 99, 0, 12, 0, 1))


def test_polygon_voxelizer_bit_exact():
    rng = np.random.default_rng(3)
    env = orc.Env(9, 64, 48)
    counts = [7, 10]
    verts = np.zeros((2, 10, 2), np.float32)
    for k, n in enumerate(counts):
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(5, 10, n)
        c = (20 + 20 * k, 24)
        verts[k, :n] = np.stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)], -1)
    ref = orc.voxelize_polygons(env, verts, counts)
    lat = Lattice(9, 64, 48, nsolid=2)
    lat.voxelize_polygons(verts, counts)
    got = lat.mask.view(48, 64).cpu().numpy()
    assert np.array_equal(got, ref) and ref.any()
