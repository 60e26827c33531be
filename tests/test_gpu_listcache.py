"""hl_step's near-body list reused across calls (hl_env.mask_epoch): results
bitwise equal to rebuilding it every call, through mask changes, other
lattices stepping on the same stream in between, and other entry points
(mass, forces) that share the stream's workspace."""
import numpy as np
import pytest
import torch

from openhomelbm_amd.lattice import Lattice
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def _lat(lattice, shape, seed, cache):
    nz = shape[0] if lattice == 27 else 1
    ny, nx = shape[-2], shape[-1]
    L = Lattice(lattice, nx, ny, nz, nsolid=2)
    L.nu = 0.02
    L.bc_type = [0, 1, 1, 1, 1, 1]
    L.bc_value = [[0.04, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
    L.set_state(random_state(lattice, shape, seed=seed).astype(np.float32))
    L._cache = cache
    return L


def _voxelize(L, caps):
    L.voxelize(np.array(caps, np.float32))
    if not L._cache:
        L._mask_epoch = 0  # 0: rebuild the list on every call


@pytest.mark.parametrize("lattice,shape", [(9, (40, 56)), (27, (16, 18, 20))])
def test_cached_list_equals_rebuilt(lattice, shape):
    z = shape[0] / 2 if lattice == 27 else 0.0
    cap_a = [[14, 20, z, 20, 22, z, 4.5, 0]]
    cap_b = [[30, 18, z, 33, 19, z, 3.5, 1], [12, 10, z, 12, 10, z, 2.5, 0]]
    runs = []
    for cache in (True, False):
        L = _lat(lattice, shape, 7, cache)
        other = _lat(lattice, shape, 8, cache)
        _voxelize(L, cap_a)
        _voxelize(other, cap_b)
        hist = torch.zeros((12, 2, 6), dtype=torch.float64, device="cuda")
        for k in range(12):
            if k == 4:
                _voxelize(L, cap_b)  # a new mask: new version
            if k == 6:
                other.step(1, force_history=hist[k:k + 1].clone())  # another lattice on the same stream
            if k == 8:
                L.mass()  # another entry point on the stream's workspace
            L.step(1, force_history=hist[k:k + 1])
        torch.cuda.synchronize()
        runs.append((L.get_state(np.float32), hist.cpu().numpy()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1]) and np.abs(runs[0][1]).max() > 0
