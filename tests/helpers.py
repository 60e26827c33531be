"""Shared helpers for the tests: seeded near-equilibrium states and metrics."""
from __future__ import annotations

import numpy as np


def random_state(lattice: int, shape, seed: int, du: float = 0.05, ds: float = 1e-3,
                 drho: float = 1e-2, dtype=np.float64) -> np.ndarray:
    """Moments near equilibrium: rho ~ 1, |u| ~ du, S ~ uu + noise.

    shape is (ny, nx) for D2Q9 and (nz, ny, nx) for D3Q27."""
    rng = np.random.default_rng(seed)
    d = 2 if lattice == 9 else 3
    rho = 1.0 + drho * rng.standard_normal(shape)
    u = du * (2 * rng.random((d,) + tuple(shape)) - 1)
    if d == 2:
        S = [u[0] * u[0], u[1] * u[1], u[0] * u[1]]
        S = [s + ds * rng.standard_normal(shape) for s in S]
        m = np.stack([rho, u[0], u[1]] + S)
    else:
        S = [u[0] * u[0], u[1] * u[1], u[2] * u[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]]
        S = [s + ds * rng.standard_normal(shape) for s in S]
        m = np.stack([rho, u[0], u[1], u[2]] + S)
    return m.astype(dtype)


def tensor_rel_l2(lattice: int, a: np.ndarray, b: np.ndarray) -> dict:
    """Relative L2 errors of a against b, per moment group.

    rho and u are compared field-wise against ||b||. S is compared as the
    tensor Pi/rho = S + I/3 (the quantity the step computes), so the
    diagonal I/3 sets the scale for every S component."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = 2 if lattice == 9 else 3
    out = {"rho": float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))}
    du = a[1:1 + d] - b[1:1 + d]
    nb = np.linalg.norm(b[1:1 + d])
    out["u"] = float(np.linalg.norm(du) / max(nb, 1e-300))
    s0 = 1 + d
    ds = a[s0:] - b[s0:]
    bt = b[s0:].copy()
    bt[:d] += 1.0 / 3.0
    out["S"] = float(np.linalg.norm(ds) / np.linalg.norm(bt))
    out["all"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return out
