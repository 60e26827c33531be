"""16-bit moment storage (FP16 / BF16 of (value - offset) * scale per field;
collision in FP32) against the FP64 oracle, BF16's stochastic rounding, and
the precision sweep driver.

Bar for one step from the same decoded input: every stored value within the
storage format's rounding (half an ulp for FP16's round-to-nearest, relative
2^-11; a whole ulp for BF16's stochastic rounding, relative 2^-7; in scaled
units of value - offset) plus twice the FP32 path's own error at that value,
plus one FP32 ulp of the field's magnitude (the two paths round
differently)."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from openhomelbm_amd import inputs, precision
from openhomelbm_amd.lattice import Lattice
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

SHAPE = (22, 20, 26)  # nz, ny, nx
ROUNDING = {"f16": 2.0 ** -11, "bf16": 2.0 ** -7}


def _lattice(dtype):
    nz, ny, nx = SHAPE
    L = Lattice(27, nx, ny, nz, dtype=dtype, nsolid=1)
    L.nu = 0.02
    L.bc_type = [0, 1, 1, 1, 2, 1]
    L.bc_value = [[0.05, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
    p = np.zeros((1, 24), np.float32)
    p[0, 0:3] = p[0, 12:15] = (nx * 0.4, ny * 0.5, nz * 0.5)
    p[0, 3:12] = p[0, 15:24] = np.eye(3).ravel()
    L.set_poses(p)
    L.voxelize(np.array([[nx * 0.35, ny * 0.5, nz * 0.5, nx * 0.45, ny * 0.5, nz * 0.5, 3.0, 0]], np.float32))
    return L, p


def _env(L, p):
    env = orc.Env(27, L.nx, L.ny, L.nz, nu=0.02, bc_type=(0, 1, 1, 1, 2, 1),
                  bc_value=((0.05, 0, 0),) + ((0, 0, 0),) * 5)
    env.mask = L.mask.cpu().numpy().reshape(SHAPE)
    env.solids = p
    return env


@pytest.mark.parametrize("uoff", [0.0, 0.05])
@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_16bit_step_within_storage_rounding(dtype, uoff):
    m0 = random_state(27, SHAPE, seed=12)
    L, p = _lattice(dtype)
    L.offset[1] = uoff
    L.set_state(m0)
    mq = L.get_state(np.float64)                 # the decoded 16-bit input
    assert np.abs(mq - m0).max() < 1e-2
    L.step(1)
    got = L.get_state(np.float64)
    env = _env(L, p)
    ref = orc.step(env, mq, "f64")
    L32, _ = _lattice("f32")
    L32.set_state(mq.astype(np.float32))
    L32.step(1)
    e32 = np.abs(L32.get_state(np.float64) - ref)
    sc = np.array(L.scale[:10]).reshape(-1, 1, 1, 1)
    stored = ref.copy()
    stored[0] -= 1.0
    stored[1] -= uoff
    # FP32 arithmetic floor: one FP32 ulp of the field's largest magnitude
    ulp32 = np.array([np.spacing(np.float32(np.abs(ref[f]).max())) for f in range(10)]).reshape(-1, 1, 1, 1)
    bound = 1.0001 * ROUNDING[dtype] * np.abs(stored * sc) / sc + 2.0 * e32 + ulp32
    err = np.abs(got - ref)
    worst = np.unravel_index(np.argmax(err - bound), err.shape)
    assert np.all(err <= bound), (worst, err[worst], bound[worst])


def test_bf16_stochastic_rounding_is_unbiased_and_deterministic():
    """A uniform fill whose scaled velocity lies 0.3 of a BF16 ulp above a
    grid point: about 30% of the cells round up, the decoded mean equals the
    value, and the same fill gives the same bits."""
    nz, ny, nx = 16, 32, 64
    L = Lattice(27, nx, ny, nz, dtype="bf16")
    sc = L.scale[1]
    lo = 0.0390625                              # 1.25 * 2^-5: a BF16 number
    ulp = 2.0 ** -5 * 2.0 ** -7                 # BF16 spacing in [2^-5, 2^-4)
    u = (lo + 0.3 * ulp) / sc                   # stored (scaled) value lo + 0.3 ulp
    L.init_uniform((u, 0.0, 0.0), set_offset=False)
    torch.cuda.synchronize()
    raw = L.fields(0)[1].float().flatten()
    up = (raw > lo).float().mean().item()
    assert set(torch.unique(raw).tolist()) == {lo, lo + ulp}
    assert abs(up - 0.3) < 0.01, up              # 32768 cells: sigma 0.0025
    mean = (raw.double() / sc).mean().item()
    assert abs(mean - u) < 0.01 * ulp / sc
    bits = L.fields(0)[1].clone()
    L.init_uniform((u, 0.0, 0.0), set_offset=False)
    assert torch.equal(bits, L.fields(0)[1])
    assert torch.equal(L.fields(0)[1], L.fields(1)[1])  # the same element index in both buffers


def test_precision_sweep_orders_errors():
    eel = inputs.Eel3D(nx=96, ny=48, nz=48, amplitude=0.4, frequency=1.5, phase_lag=0.8, length=40.0, radius=4.0)
    eel.time_scale = min(1.0, inputs.EEL_VMAX / eel.max_speed())
    t = precision.sweep(steps=40, eel=eel)
    e = t["errors"]
    assert set(e) == {"f64", "f32", "f16", "bf16"}
    assert e["f32"]["u_rel_l2"] < 1e-5
    assert e["f32"]["u_rel_l2"] < min(e["f16"]["u_rel_l2"], e["bf16"]["u_rel_l2"])
    assert max(e["f16"]["u_rel_l2"], e["bf16"]["u_rel_l2"]) < 2e-3
    assert max(e["f16"]["mean_drag_rel_err"], e["bf16"]["mean_drag_rel_err"]) < 5e-3
    assert e["f32"]["mean_drag_rel_err"] < 1e-4
    assert np.isfinite(e["f64"]["mean_drag"]) and e["f64"]["mean_drag"] != 0.0
