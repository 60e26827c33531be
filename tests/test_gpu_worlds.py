"""Several D3Q27 lattices stepped by one library call (hl_step_worlds, the
world dimension of LBM_Solver3D.step, envs/lbm3d/lbm_solver_3d.py:104-108).

The tiled pass of each world in the shared launch runs the same per-cell
arithmetic as that world's own launch, its near-body cells and force rows
are its own, so every world must end bitwise where its own hl_step call
ends: states and force histories are compared for exact equality."""
import ctypes as C

import numpy as np
import pytest
import torch

from openhomelbm_amd import _native as N
from openhomelbm_amd.lattice import Lattice, step_worlds
from openhomelbm_amd.lbm3d import LBM_Solver3D
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

# (nx, ny, nz); every size fits the 64x8 tile (nx % 64 and ny % 8 not 1 or 2)
SIZES = [(40, 24, 20), (72, 20, 18), (64, 16, 16), (48, 28, 23), (35, 19, 25), (96, 40, 30), (20, 12, 14),
         (56, 22, 21)]


@pytest.fixture(autouse=True)
def _restore_tiling():
    saved = N.get_tiling(27)
    yield
    N.set_tiling(27, *saved)


def _world(k, dtype="f32"):
    nx, ny, nz = SIZES[k]
    rng = np.random.default_rng(100 + k)
    body = k % 3 != 2  # every third world has no body
    L = Lattice(27, nx, ny, nz, dtype=dtype, nsolid=2 if body else 0)
    L.nu = 0.01 + 0.002 * k
    L.bc_type = [0, 1, 2, 2, 1, 2] if k % 2 else [0, 2, 1, 1, 2, 1]
    L.bc_value = [[0.04 + 0.001 * k, 0.0, 0.002]] + [[0.0, 0.0, 0.0]] * 5
    caps = None
    if body:
        r = 0.12 * min(nx, ny, nz)
        c0 = (0.35 * nx, 0.5 * ny, 0.5 * nz)
        c1 = (0.6 * nx, 0.45 * ny, 0.55 * nz)
        caps = np.array([[*c0, *c0, r, 0], [c1[0] - 2, c1[1], c1[2], c1[0] + 2, c1[1] + 1, c1[2], 0.8 * r, 1]],
                        np.float32)
        poses = np.zeros((2, 24), np.float32)
        poses[:, 3:12] = poses[:, 15:24] = np.eye(3).ravel()
        poses[0, 0:3] = poses[0, 12:15] = c0
        poses[1, 0:3] = c1
        poses[1, 12:15] = (c1[0] - 0.02, c1[1] + 0.01, c1[2])
        L.set_poses(poses)
        L.voxelize(caps)
    if k == 3:  # one world with a body-force field (a chunk of its own)
        L.set_force(2e-4 * rng.standard_normal((3, nz, ny, nx)).astype(np.float32))
    L.set_state(random_state(27, (nz, ny, nx), seed=200 + k).astype(np.float32))
    return L


def _history(L, nsteps):
    if L.nsolid == 0:
        return None
    return torch.full((nsteps, L.nsolid, 6), float("nan"), dtype=torch.float64, device=L.device)


def _fits(L, tile):
    tx, ty, _ = tile
    return L.nx % tx not in (1, 2) and L.ny % ty not in (1, 2)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_step_worlds_equal_per_world_calls(dtype):
    """Eight worlds of eight sizes (chunks of six and two, one world with a
    body-force field that takes a chunk of its own), bodies in most, force
    histories in those: after three steps every state and force history is
    bitwise equal to the world's own hl_step calls."""
    tile = (64, 8, 16)
    N.set_tiling(27, *tile)
    nsteps = 3
    many = [_world(k, dtype) for k in range(len(SIZES))]
    alone = [_world(k, dtype) for k in range(len(SIZES))]
    assert all(_fits(L, tile) for L in many)  # the shared tiled launches run
    hm = [_history(L, nsteps) for L in many]
    ha = [_history(L, nsteps) for L in alone]
    envs = (N.HLEnv * len(many))(*[L._step_env() for L in many])
    ptrs = (C.c_void_p * len(many))(*[h.data_ptr() if h is not None else None for h in hm])
    N.check(many[0].lib.hl_step_worlds(envs, len(many), 0, nsteps, N.HL_KERNEL_AUTO, ptrs, None), "hl_step_worlds")
    for L in many:
        L.cur ^= nsteps & 1
    for L, h in zip(alone, ha):
        L.step(nsteps, N.HL_KERNEL_AUTO, force_history=h, _tuned=False)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(many, alone)):
        assert np.array_equal(a.get_state(np.float64), b.get_state(np.float64)), f"world {k} state"
        if hm[k] is not None:
            assert torch.equal(hm[k], ha[k]), f"world {k} forces"
            assert torch.isfinite(hm[k]).all()


def test_step_worlds_python_and_solver():
    """step_worlds (the Python call) and LBM_Solver3D with three worlds with a
    capsule each, stepped one step at a time: each world equals a solver of
    its own, state and forces."""
    nx, ny, nz = 48, 24, 20
    multi = LBM_Solver3D(nx, ny, nz, solid_num=1, nworld=3)
    solo = [LBM_Solver3D(nx, ny, nz, solid_num=1, nworld=1) for _ in range(3)]
    for k in range(3):
        for s in (multi.flows[k], solo[k].flows[0]):
            s.set_capsule(0, (-3.0, 0.0, 0.0), (3.0, 0.0, 0.0), 3.0 + 0.5 * k)
            s.solid_position[0] = (18.0 + k, 12.0, 10.0)
            m = random_state(27, (nz, ny, nx), seed=300 + k).astype(np.float32)
            s.lattice.set_state(m)
    for step in range(4):
        for k in range(3):
            for s in (multi.flows[k], solo[k].flows[0]):
                s.solid_position[0] = (18.0 + k + 0.1 * step, 12.0, 10.0)
        multi.step()
        for sv in solo:
            sv.step()
    torch.cuda.synchronize()
    for k in range(3):
        a, b = multi.flows[k], solo[k].flows[0]
        assert np.array_equal(a.lattice.get_state(np.float64), b.lattice.get_state(np.float64)), k
        assert torch.equal(a._forces, b._forces), k


def test_step_worlds_errors_name_the_world():
    a = Lattice(27, 40, 24, 20, nsolid=0)
    b = Lattice(27, 40, 24, 20, dtype="bf16", nsolid=0)
    with pytest.raises(ValueError):
        step_worlds([a, b], 1)
    envs = (N.HLEnv * 2)(a._step_env(), a._step_env())
    envs[1].nx = 2  # every extent must be >= 3
    rc = a.lib.hl_step_worlds(envs, 2, 0, 1, N.HL_KERNEL_AUTO, None, None)
    assert rc < 0
    assert b"world 1" in a.lib.hl_last_error()
