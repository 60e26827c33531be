"""Seeded synthetic workloads (NumPy PCG64; FP64 draws cast to FP32).

The same builders feed the GPU runs, the CPU oracle baseline and the tests,
so every consumer sees identical arrays.

* d2q9_karman: D2Q9 512x128 Karman flow, disc of diameter 16 at (96, 65),
  inlet u = 0.05, nu = 0.008 (tau = 0.524), noise 1e-3, seed 1001.
* d3q27_sphere: D3Q27 256x128x128 flow past a static sphere of diameter 24
  at (64, 64, 65), inlet u = 0.05, Re = 200 (nu = 0.006), noise 1e-3, seed 2002.

Initial state: the reference's uniform-flow initialisation
(set_uniform_flow_kernel, envs/lbm/karman/karman_lbm_env.py:15-53 and its 3D
twin) at the inlet velocity, then seeded uniform noise of the stated
amplitude added to every velocity component.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class Workload:
    name: str
    lattice: int
    nx: int
    ny: int
    nz: int
    nu: float
    u_in: Tuple[float, float, float]
    bc_type: Tuple[int, ...]
    bc_value: Tuple[Tuple[float, float, float], ...]
    noise: float
    seed: int
    capsules: np.ndarray            # [n, 8] ax, ay, az, bx, by, bz, r, solid_id
    poses: np.ndarray               # [nsolid, 8 | 24] (static: current == last)
    steps: int = 1000
    description: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def shape(self):
        return (self.ny, self.nx) if self.lattice == 9 else (self.nz, self.ny, self.nx)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def velocity_noise(self) -> np.ndarray:
        """[d, (z,) y, x] FP32 noise added to u after the uniform init."""
        d = 2 if self.lattice == 9 else 3
        rng = np.random.Generator(np.random.PCG64(self.seed))
        return (self.noise * (2.0 * rng.random((d,) + self.shape) - 1.0)).astype(np.float32)


def _static_pose_2d(x, y):
    return np.array([[x, y, 1, 0, x, y, 1, 0]], np.float32)


def _static_pose_3d(x, y, z):
    p = np.zeros((1, 24), np.float32)
    p[0, 0:3] = (x, y, z)
    p[0, 12:15] = (x, y, z)
    p[0, 3:12] = np.eye(3).ravel()
    p[0, 15:24] = np.eye(3).ravel()
    return p


def d2q9_karman(seed: int = 1001) -> Workload:
    """512x128, disc D = 16 at (96, 65), Re = 100 (nu = 0.008). The reference
    has no free-slip face type; top and bottom use its zero-gradient type 1
    as in configs/realtime_2d/karman2d.json:54 (bc_type [0, 1, 1, 1])."""
    u = 0.05
    return Workload(
        name="d2q9_karman_512x128", lattice=9, nx=512, ny=128, nz=1, nu=0.008, u_in=(u, 0.0, 0.0),
        bc_type=(0, 1, 1, 1), bc_value=((u, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5, noise=1e-3, seed=seed,
        capsules=np.array([[96, 65, 0, 96, 65, 0, 8, 0]], np.float32), poses=_static_pose_2d(96, 65),
        steps=2000, description="D2Q9 Karman flow, 512x128, disc D=16 at (96,65), u=0.05, Re=100")


def d3q27_sphere(seed: int = 2002, nx: int = 256, ny: int = 128, nz: int = 128) -> Workload:
    """256x128x128, static sphere D = 24 at (64, 64, 65), Re = 200 with
    U = 0.05 and L = D (nu = 0.006); inlet at x = 0, zero-gradient elsewhere."""
    u = 0.05
    D = 24.0
    return Workload(
        name=f"d3q27_sphere_{nx}x{ny}x{nz}", lattice=27, nx=nx, ny=ny, nz=nz, nu=u * D / 200.0,
        u_in=(u, 0.0, 0.0), bc_type=(0, 1, 1, 1, 1, 1), bc_value=((u, 0.0, 0.0),) + ((0.0, 0.0, 0.0),) * 5,
        noise=1e-3, seed=seed,
        capsules=np.array([[64, 64, 65, 64, 64, 65, D / 2, 0]], np.float32), poses=_static_pose_3d(64, 64, 65),
        steps=1000, description=f"D3Q27 flow past a static sphere D=24 at (64,64,65), {nx}x{ny}x{nz}, Re=200")


def build_lattice(w: Workload, dtype: str = "f32"):
    """GPU Lattice initialised with the workload's state, bodies and faces."""
    import torch

    from .lattice import Lattice

    lat = Lattice(w.lattice, w.nx, w.ny, w.nz, dtype=dtype, nsolid=w.poses.shape[0])
    lat.nu = w.nu
    lat.bc_type = list(w.bc_type) + [1] * (6 - len(w.bc_type))
    lat.bc_value = [list(v) for v in w.bc_value]
    lat.set_poses(w.poses)
    lat.voxelize(w.capsules)
    lat.init_uniform(w.u_in)
    noise = torch.from_numpy(w.velocity_noise()).to(lat.device)
    d = noise.shape[0]
    for b in (0, 1):
        f = lat.fields(b)
        if dtype in ("f16", "bf16"):
            cur = f[1:1 + d].float() / lat.scale[1]
            f[1:1 + d].copy_(((cur + noise) * lat.scale[1]).to(f.dtype))
        else:
            f[1:1 + d] += noise.to(f.dtype)
    return lat
