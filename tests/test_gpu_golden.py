"""The HIP path against the frozen golden fixtures (tests/golden/): the gather
kernel bit for bit in FP32 and FP64, the default (tiled) path within the
tiled bar, forces within FP64 summation-order rounding."""
import numpy as np
import pytest
import torch

from openhomelbm_amd import _native as N
from openhomelbm_amd.lattice import Lattice
from tests.helpers import tensor_rel_l2
from tests.test_golden import NAMES, load

pytestmark = pytest.mark.gpu


def _lattice(g, env, dtype):
    L = Lattice(env.lattice, env.nx, env.ny, env.nz, dtype=dtype, nsolid=g["poses"].shape[0])
    L.nu = float(g["nu"])
    L.bc_type = [int(t) for t in g["bc_type"]]
    L.bc_value = [[float(v) for v in r] for r in g["bc_value"]]
    L.set_poses(g["poses"])
    L.voxelize(g["caps"])
    L.set_force(g["force"])
    L.set_state(g["m0"].astype(np.float64))
    return L


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_gather_reproduces_golden(name, dtype):
    g, env = load(name)
    L = _lattice(g, env, dtype)
    assert np.array_equal(L.mask.cpu().numpy().reshape(g["mask"].shape), g["mask"])
    h = torch.zeros((3, L.nsolid, 6), dtype=torch.float64, device=L.device)
    L.step(3, N.HL_KERNEL_GATHER, force_history=h)
    got = L.get_state(np.float32 if dtype == "f32" else np.float64)
    assert np.array_equal(got, g[f"{dtype}_step3"])
    np.testing.assert_allclose(h.cpu().numpy(), g[f"{dtype}_forces"], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("name", NAMES)
def test_default_path_within_bar(name):
    g, env = load(name)
    L = _lattice(g, env, "f32")
    for k in range(3):
        L.step(1)
        err = tensor_rel_l2(env.lattice, L.get_state(np.float64), g[f"f32_step{k + 1}"].astype(np.float64))
        assert err["all"] <= 1e-6, (k, err)
