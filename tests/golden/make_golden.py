"""Writes the golden fixtures of this directory from the CPU oracle.

The reference ships no golden vectors and its Warp kernels cannot run here, so
these fixtures are the oracle's own outputs on small seeded inputs, frozen.
They pin the oracle against regressions (tests/test_golden.py, CPU) and give
the GPU path fixed expected outputs (tests/test_gpu_golden.py).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as orc  # noqa: E402
from tests.helpers import random_state  # noqa: E402

CASES = {
    # lattice, (nz,) ny, nx, nu, bc_type, capsules, seed
    "d2q9_16x12": (9, (12, 16), 0.03, (0, 2, 1, 1),
                   [[5.0, 6.0, 0, 8.5, 6.5, 0, 2.2, 0], [11.5, 4.0, 0, 11.5, 4.0, 0, 1.6, 1]], 101),
    "d3q27_12x10x8": (27, (8, 10, 12), 0.04, (0, 1, 2, 1, 1, 2),
                      [[4.5, 5.0, 4.0, 6.5, 5.0, 4.0, 1.8, 0], [9.0, 4.0, 3.5, 9.0, 4.0, 3.5, 1.4, 1]], 202),
}
STEPS = 3


def poses_for(lattice, caps):
    """Pose tables of the two solids: solid 0 translates and rotates, solid 1 is static."""
    n = len(caps)
    if lattice == 9:
        p = np.zeros((n, 8), np.float32)
        for k, c in enumerate(caps):
            p[k] = [c[0], c[1], 1, 0, c[0], c[1], 1, 0]
        a = 0.03
        p[0, 2:4] = [np.cos(a), np.sin(a)]
        p[0, 4:6] = [caps[0][0] - 0.02, caps[0][1] + 0.01]
        return p
    p = np.zeros((n, 24), np.float32)
    for k, c in enumerate(caps):
        p[k, 0:3] = p[k, 12:15] = c[0:3]
        p[k, 3:12] = p[k, 15:24] = np.eye(3).ravel()
    a = 0.03
    p[0, 3:12] = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]).ravel()
    p[0, 12:15] = np.asarray(caps[0][0:3]) - [0.02, -0.01, 0.005]
    return p


def build(name):
    lattice, shape, nu, bc_type, caps, seed = CASES[name]
    nx, ny = shape[-1], shape[-2]
    nz = shape[0] if lattice == 27 else 1
    caps = np.array(caps, np.float32)
    bc_value = ((0.04, 0.005, 0.0),) + ((0.0, 0.0, 0.0),) * 5
    env = orc.Env(lattice, nx, ny, nz, nu=nu, bc_type=bc_type, bc_value=bc_value)
    env.mask = orc.voxelize(env, caps)
    env.solids = poses_for(lattice, caps)
    d = 2 if lattice == 9 else 3
    env.force = (1e-4 * np.random.default_rng(seed + 1).standard_normal((d,) + shape)).astype(np.float32)
    m0 = random_state(lattice, shape, seed=seed)
    out = {"m0": m0, "caps": caps, "poses": env.solids, "force": env.force,
           "mask": env.mask, "nu": np.float32(nu), "bc_type": np.array(bc_type, np.int32),
           "bc_value": np.array(bc_value, np.float32)}
    for prec, dt in (("f32", np.float32), ("f64", np.float64)):
        m = m0.astype(dt)
        forces = []
        for k in range(STEPS):
            forces.append(orc.compute_forces(env, m, prec))
            m = orc.step(env, m, prec)
            out[f"{prec}_step{k + 1}"] = m
        out[f"{prec}_forces"] = np.array(forces)
    return out


def main():
    for name in CASES:
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **build(name))
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
