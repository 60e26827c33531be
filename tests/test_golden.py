"""The CPU oracle reproduces the frozen golden fixtures (tests/golden/) bit for
bit: masks, FP32 and FP64 states after 1-3 steps, and per-step forces."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["d2q9_16x12", "d3q27_12x10x8"]


def load(name):
    g = dict(np.load(os.path.join(GOLD, f"{name}.npz")))
    lattice = 9 if name.startswith("d2q9") else 27
    shape = g["m0"].shape[1:]
    nx, ny = shape[-1], shape[-2]
    nz = shape[0] if lattice == 27 else 1
    env = orc.Env(lattice, nx, ny, nz, nu=float(g["nu"]), bc_type=tuple(int(t) for t in g["bc_type"]),
                  bc_value=tuple(tuple(float(v) for v in r) for r in g["bc_value"]))
    return g, env


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g, env = load(name)
    mask = orc.voxelize(env, g["caps"])
    assert np.array_equal(mask, g["mask"])
    env.mask = mask
    env.solids = g["poses"]
    env.force = g["force"]
    for prec, dt in (("f32", np.float32), ("f64", np.float64)):
        m = g["m0"].astype(dt)
        for k in range(3):
            f = orc.compute_forces(env, m, prec)
            assert np.array_equal(f, g[f"{prec}_forces"][k])
            m = orc.step(env, m, prec)
            assert np.array_equal(m, g[f"{prec}_step{k + 1}"]), (prec, k)
    assert np.abs(g["f32_forces"]).max() > 0
