"""The allocation-free step entry point (hl_step_noalloc) against hl_step:
the same states and force histories, bitwise, over consecutive calls whose
lattices shrink, and a clear error when the near-body work exceeds its
fixed scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

from openhomelbm_amd import _native as N
from openhomelbm_amd.lattice import Lattice
from tests.helpers import random_state
from tests.test_gpu_parity import _gpu_lattice
from tests.test_gpu_tiled import _tiled_scene

pytestmark = pytest.mark.gpu


def _noalloc_step(lat, nsteps, hist):
    e = lat._step_env()
    rc = lat.lib.hl_step_noalloc(C.byref(e), lat.cur, int(nsteps), N.HL_KERNEL_AUTO, hist.data_ptr(),
                                 torch.cuda.current_stream().cuda_stream)
    N.check(rc, "hl_step_noalloc")
    lat.cur ^= nsteps & 1


def _small_scene(seed):
    """A 16x14x12 D3Q27 channel with one capsule (smaller than _tiled_scene)."""
    from oracle import oracle as orc
    nz, ny, nx = 12, 14, 16
    caps = np.array([[7, 7, 6, 9, 7, 6, 2.0, 0]], np.float32)
    poses = np.zeros((1, 24), np.float32)
    poses[:, 3:12] = poses[:, 15:24] = np.eye(3).ravel()
    poses[0, 0:3] = poses[0, 12:15] = (8, 7, 6)
    env = orc.Env(27, nx, ny, nz, nu=0.03, bc_type=(0, 1, 1, 1, 1, 1),
                  bc_value=((0.04, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5)
    return env, caps, poses, None


@pytest.mark.parametrize("scene,shape,nsteps", [(_tiled_scene, (34, 40, 72), 3), (_small_scene, (12, 14, 16), 2)])
def test_noalloc_equals_step(scene, shape, nsteps):
    env, caps, poses, force = scene(21)
    m0 = random_state(27, shape, seed=23).astype(np.float32)
    a = _gpu_lattice(env, caps, poses, force)
    b = _gpu_lattice(env, caps, poses, force)
    for L in (a, b):
        L.set_state(m0)
    ha = torch.zeros((nsteps, a.nsolid, 6), dtype=torch.float64, device=a.device)
    hb = torch.full_like(ha, np.nan)
    a.step(nsteps, force_history=ha)
    _noalloc_step(b, nsteps, hb)
    torch.cuda.synchronize()
    assert a.cur == b.cur
    assert np.array_equal(a.get_state(np.float32), b.get_state(np.float32))
    assert torch.equal(ha, hb)
    assert float(ha.abs().max()) > 0


def test_consecutive_calls_with_shrinking_input():
    """Two calls in a row on one GPU share the scratch: the first on the
    72x40x34 scene with its two bodies, the second on a smaller lattice;
    each equals hl_step bitwise."""
    outs = []
    for scene, shape, nsteps in ((_tiled_scene, (34, 40, 72), 2), (_small_scene, (12, 14, 16), 3)):
        env, caps, poses, force = scene(5)
        m0 = random_state(27, shape, seed=29).astype(np.float32)
        a = _gpu_lattice(env, caps, poses, force)
        b = _gpu_lattice(env, caps, poses, force)
        for L in (a, b):
            L.set_state(m0)
        ha = torch.zeros((nsteps, a.nsolid, 6), dtype=torch.float64, device=a.device)
        hb = torch.zeros_like(ha)
        _noalloc_step(b, nsteps, hb)
        a.step(nsteps, force_history=ha)
        torch.cuda.synchronize()
        outs.append((np.array_equal(a.get_state(np.float32), b.get_state(np.float32)), torch.equal(ha, hb)))
    assert outs == [(True, True), (True, True)]


def test_d2q9_gather_with_forces():
    from tests.test_gpu_parity import _scene
    env, caps, poses, force = _scene(9, (40, 56), 3)
    m0 = random_state(9, (40, 56), seed=31).astype(np.float32)
    a = _gpu_lattice(env, caps, poses, force)
    b = _gpu_lattice(env, caps, poses, force)
    for L in (a, b):
        L.set_state(m0)
    ha = torch.zeros((4, a.nsolid, 6), dtype=torch.float64, device=a.device)
    hb = torch.zeros_like(ha)
    a.step(4, force_history=ha)
    _noalloc_step(b, 4, hb)
    torch.cuda.synchronize()
    assert np.array_equal(a.get_state(np.float32), b.get_state(np.float32))
    assert torch.equal(ha, hb)


def test_oversized_near_body_work_is_rejected():
    """A mask box of 128^3 cells needs more than the fixed scratch: the call
    fails with HL_E_NOMEM before any launch, and the state is untouched."""
    lat = Lattice(27, 128, 128, 128, nsolid=1)
    lat.init_uniform((0.05, 0.0, 0.0))
    mask = np.zeros((128, 128, 128), np.uint8)
    mask[1, 1, 1] = mask[126, 126, 126] = 1
    lat.set_mask(mask)
    before = lat.fields().clone()
    h = torch.zeros((1, 1, 6), dtype=torch.float64, device=lat.device)
    e = lat._step_env()
    rc = lat.lib.hl_step_noalloc(C.byref(e), lat.cur, 1, N.HL_KERNEL_AUTO, h.data_ptr(),
                                 torch.cuda.current_stream().cuda_stream)
    assert rc == -4
    assert "scratch" in lat.lib.hl_last_error().decode()
    torch.cuda.synchronize()
    assert torch.equal(before, lat.fields())
