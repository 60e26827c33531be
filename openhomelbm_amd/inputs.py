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
        for k in range(d):
            f[1 + k].copy_(lat.encode(1 + k, lat.decode(1 + k, f[1 + k]) + noise[k]))
    return lat


# ----------------------------------------------------------------- eels

EEL_SIZES = ((192, 96), (256, 128), (384, 160), (512, 192))
EEL_SEGMENTS = 8
EEL_RADIUS = 4.0
EEL_U = 0.05
EEL_VMAX = 0.1


@dataclass
class Eel:
    """An 8-capsule planar chain with sinusoidal CPG joint angles.

    Joint j (1..7, between segments j-1 and j) follows
        theta_j(t) = A * env_j * sin(omega * t - j * phi),  env_j = 0.05 + 0.95 * j / 7,
    the envelope of the reference's eel_wave preset
    (tools/lbm2d_realtime_control.py:337-370); segment 0 keeps heading 0 (the
    chain trails downstream of its head). omega = 2 pi f / 1000 per lattice
    step, slowed (time_scale < 1) so that no body point moves faster than
    EEL_VMAX lattice units per step: taken literally, in-phase chains reach
    tail speeds above 1 (faster than the lattice sound speed).
    """

    nx: int
    ny: int
    amplitude: float
    frequency: float   # cycles per 1000 steps, before time_scale
    phase_lag: float
    head: Tuple[float, float]
    seg_len: float
    radius: float = EEL_RADIUS
    time_scale: float = 1.0

    @property
    def length(self) -> float:
        return self.seg_len * EEL_SEGMENTS

    @property
    def nu(self) -> float:
        """Re = 300 on the chain length at U = 0.05."""
        return EEL_U * self.length / 300.0

    def joints(self, t) -> np.ndarray:
        """FP64 joint positions [..., 9, 2] at (possibly fractional) times t."""
        t = np.asarray(t, dtype=np.float64)[..., None]
        j = np.arange(EEL_SEGMENTS, dtype=np.float64)
        env = 0.05 + 0.95 * j / (EEL_SEGMENTS - 1)
        omega = 2.0 * np.pi * self.frequency / 1000.0 * self.time_scale
        theta = self.amplitude * env * np.sin(omega * t - j * self.phase_lag)
        theta[..., 0] = 0.0
        heading = np.cumsum(theta, axis=-1)
        seg = self.seg_len * np.stack([np.cos(heading), np.sin(heading)], axis=-1)
        p = np.zeros(t.shape[:-1] + (EEL_SEGMENTS + 1, 2))
        p[..., 0, :] = self.head
        p[..., 1:, :] = self.head + np.cumsum(seg, axis=-2)
        return p

    def tables(self, t0: int, count: int):
        """FP32 capsule [count, 8, 8] and pose [count, 8, 8] tables for steps
        t0 .. t0+count-1 (computed in FP64, rounded once). Capsule j runs from
        joint j to joint j+1 with solid id j; pose j is (joint j, cos, sin of
        the segment heading) at t and at t - 1."""
        t = np.arange(t0, t0 + count, dtype=np.float64)
        now, last = self.joints(t), self.joints(t - 1.0)
        caps = np.zeros((count, EEL_SEGMENTS, 8))
        caps[..., 0:2] = now[:, :-1]
        caps[..., 3:5] = now[:, 1:]
        caps[..., 6] = self.radius
        caps[..., 7] = np.arange(EEL_SEGMENTS)
        poses = np.zeros((count, EEL_SEGMENTS, 8))
        for k, p in ((0, now), (4, last)):
            d = p[:, 1:] - p[:, :-1]
            h = np.arctan2(d[..., 1], d[..., 0])
            poses[..., k:k + 2] = p[:, :-1]
            poses[..., k + 2] = np.cos(h)
            poses[..., k + 3] = np.sin(h)
        return caps.astype(np.float32), poses.astype(np.float32)

    def box(self, t0: int, count: int):
        """Half-open cell box covering the chain over the steps, padded by 1."""
        p = self.joints(np.arange(t0, t0 + count, dtype=np.float64))
        lo = p.reshape(-1, 2).min(0) - self.radius
        hi = p.reshape(-1, 2).max(0) + self.radius
        x0, y0 = int(np.floor(lo[0])) - 1, int(np.floor(lo[1])) - 1
        x1, y1 = int(np.ceil(hi[0])) + 2, int(np.ceil(hi[1])) + 2
        return (max(0, x0), min(self.nx, x1), max(0, y0), min(self.ny, y1), 0, 1)

    def max_speed(self, period_samples: int = 512) -> float:
        """Largest joint displacement per lattice step over one CPG period."""
        omega = 2.0 * np.pi * self.frequency / 1000.0 * self.time_scale
        period = 2.0 * np.pi / omega
        t = np.linspace(0.0, period, period_samples)
        return float(np.abs(self.joints(t + 1.0) - self.joints(t)).max())


def eel_batch(n: int = 64, seed: int = 3003) -> List[Eel]:
    """n eels with seeded grid sizes and CPG parameters: size from
    {192x96, 256x128, 384x160, 512x192}, chain length 0.3 nx, radius 4,
    A ~ U[0.2, 0.6] rad, f ~ U[0.5, 2] per 1000 steps, phi ~ U[0, pi/2].
    Head at (0.25 nx, ny / 2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    eels = []
    for _ in range(n):
        nx, ny = EEL_SIZES[int(rng.integers(0, len(EEL_SIZES)))]
        A = float(rng.uniform(0.2, 0.6))
        f = float(rng.uniform(0.5, 2.0))
        phi = float(rng.uniform(0.0, np.pi / 2))
        e = Eel(nx=nx, ny=ny, amplitude=A, frequency=f, phase_lag=phi, head=(0.25 * nx, 0.5 * ny),
                seg_len=0.3 * nx / EEL_SEGMENTS)
        v = e.max_speed()
        if v > EEL_VMAX:
            e.time_scale = EEL_VMAX / v
        eels.append(e)
    return eels


def eel_state(e: Eel, index: int, seed: int = 3003) -> np.ndarray:
    """FP32 initial moments [6, ny, nx]: rho = 1, u = (0.05, 0) + 1e-3 noise
    (seeded per environment), S = 0, the reference's uniform-flow values."""
    rng = np.random.Generator(np.random.PCG64([seed, index]))
    m = np.zeros((6, e.ny, e.nx), np.float64)
    m[0] = 1.0
    m[1] = EEL_U
    m[1:3] += 1e-3 * (2.0 * rng.random((2, e.ny, e.nx)) - 1.0)
    return m.astype(np.float32)


def build_eel_batch(n: int = 64, seed: int = 3003, steps: int = 2000):
    """GPU Batch of n eel environments with kinematics tables for `steps` steps.

    Faces: velocity inlet (0.05, 0) on the left, zero-gradient outlet, free-slip
    top and bottom. Returns (batch, lattices, eels)."""
    from .batch import Batch
    from .lattice import Lattice

    eels = eel_batch(n, seed)
    lats = []
    for i, e in enumerate(eels):
        L = Lattice(9, e.nx, e.ny, nsolid=EEL_SEGMENTS)
        L.nu = e.nu
        L.bc_type = [0, 2, 1, 2, 1, 1]
        L.bc_value = [[EEL_U, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
        L.set_state(eel_state(e, i, seed))
        lats.append(L)
    b = Batch(lats)
    for i, e in enumerate(eels):
        caps, poses = e.tables(0, steps)
        b.set_kinematics(i, caps, poses, e.box(0, steps))
    return b, lats, eels


def eel_alone(e: Eel, index: int, seed: int = 3003, steps: int = 2000, dtype: str = "f64"):
    """Environment `index` of an eel batch stepped on its own (one Lattice,
    the gather kernel, masks re-voxelized every step from the same tables):
    the reference computation of the batched run, in FP64 by default.
    Returns the device force history [steps, 8, 6] (FP64)."""
    import torch

    from . import _native as N
    from .lattice import Lattice

    L = Lattice(9, e.nx, e.ny, dtype=dtype, nsolid=EEL_SEGMENTS)
    L.nu = e.nu
    L.bc_type = [0, 2, 1, 2, 1, 1]
    L.bc_value = [[EEL_U, 0.0, 0.0]] + [[0.0, 0.0, 0.0]] * 5
    L.set_state(eel_state(e, index, seed))
    caps, poses = e.tables(0, steps)
    box = e.box(0, steps)
    caps_d, poses_d = torch.from_numpy(caps).to(L.device), torch.from_numpy(poses).to(L.device)
    hist = torch.zeros((steps, EEL_SEGMENTS, 6), dtype=torch.float64, device=L.device)
    for t in range(steps):
        L.voxelize_device(caps_d[t], box)
        L.solids[:EEL_SEGMENTS].copy_(poses_d[t])
        L.step(1, N.HL_KERNEL_GATHER, force_history=hist[t:t + 1])
    return hist


def batched_drag_check(n: int = 64, seed: int = 3003, steps: int = 400):
    """Mean drag of every eel of the batched FP32 run against the same eel
    stepped alone in FP64 over the same steps: per-eel relative errors of the
    mean total drag (sum of Fx over the chain), their maximum and the error of
    the batch-wide mean."""
    import torch

    b, lats, eels = build_eel_batch(n, seed, steps=steps)
    h = b.history_buffer(steps)
    b.step(steps, force_history=h)
    d32 = h[:, :, :EEL_SEGMENTS, 0].sum(-1).mean(0).cpu().numpy()  # [n]
    b.close()
    del b, lats
    d64 = np.array([float(eel_alone(e, i, seed, steps)[:, :, 0].sum(-1).mean()) for i, e in enumerate(eels)])
    torch.cuda.synchronize()
    rel = np.abs(d32 - d64) / np.abs(d64)
    return {"envs": n, "steps": steps, "mean_drag_rel_err_max": float(rel.max()),
            "mean_drag_rel_err_median": float(np.median(rel)),
            "batch_mean_drag_rel_err": float(abs(d32.mean() - d64.mean()) / abs(d64.mean())),
            "reference": "each eel alone, FP64 gather kernel, same kinematics tables"}


@dataclass
class Eel3D:
    """Articulated capsule chain swimming in the x-y plane of a D3Q27 lattice
    (the precision-sweep body): nseg capsules of equal length, joint angles
    theta_j = A env_j sin(omega t - j phi) as Eel, time slowed so that no body
    point moves faster than EEL_VMAX per step."""

    nx: int
    ny: int
    nz: int
    amplitude: float
    frequency: float
    phase_lag: float
    length: float = 160.0
    radius: float = 8.0
    nseg: int = 10
    time_scale: float = 1.0

    @property
    def head(self):
        return (0.25 * self.nx, 0.5 * self.ny, 0.5 * self.nz)

    @property
    def seg_len(self) -> float:
        return self.length / self.nseg

    def nu(self, re: float = 500.0) -> float:
        return EEL_U * self.length / re

    def joints(self, t) -> np.ndarray:
        t = np.asarray(t, dtype=np.float64)[..., None]
        j = np.arange(self.nseg, dtype=np.float64)
        env = 0.05 + 0.95 * j / (self.nseg - 1)
        omega = 2.0 * np.pi * self.frequency / 1000.0 * self.time_scale
        theta = self.amplitude * env * np.sin(omega * t - j * self.phase_lag)
        theta[..., 0] = 0.0
        heading = np.cumsum(theta, axis=-1)
        seg = self.seg_len * np.stack([np.cos(heading), np.sin(heading)], axis=-1)
        p = np.zeros(t.shape[:-1] + (self.nseg + 1, 2))
        p[..., 0, :] = self.head[:2]
        p[..., 1:, :] = np.asarray(self.head[:2]) + np.cumsum(seg, axis=-2)
        return p

    def max_speed(self, samples: int = 512) -> float:
        omega = 2.0 * np.pi * self.frequency / 1000.0 * self.time_scale
        t = np.linspace(0.0, 2.0 * np.pi / omega, samples)
        return float(np.abs(self.joints(t + 1.0) - self.joints(t)).max())

    def tables(self, t0: int, count: int):
        """FP32 capsule [count, nseg, 8] and 3D pose [count, nseg, 24] tables:
        capsule j from joint j to joint j+1 at z = nz/2, solid id j; pose j =
        (joint j, rotation about z by the segment heading) at t and t - 1."""
        t = np.arange(t0, t0 + count, dtype=np.float64)
        now, last = self.joints(t), self.joints(t - 1.0)
        zc = self.head[2]
        caps = np.zeros((count, self.nseg, 8))
        caps[..., 0:2] = now[:, :-1]
        caps[..., 2] = zc
        caps[..., 3:5] = now[:, 1:]
        caps[..., 5] = zc
        caps[..., 6] = self.radius
        caps[..., 7] = np.arange(self.nseg)
        poses = np.zeros((count, self.nseg, 24))
        for k, p in ((0, now), (12, last)):
            d = p[:, 1:] - p[:, :-1]
            h = np.arctan2(d[..., 1], d[..., 0])
            c, s = np.cos(h), np.sin(h)
            poses[..., k + 0:k + 2] = p[:, :-1]
            poses[..., k + 2] = zc
            R = poses[..., k + 3:k + 12]
            R[..., 0], R[..., 1], R[..., 3], R[..., 4], R[..., 8] = c, -s, s, c, 1.0
        return caps.astype(np.float32), poses.astype(np.float32)

    def box(self, t0: int, count: int):
        p = self.joints(np.arange(t0, t0 + count, dtype=np.float64)).reshape(-1, 2)
        lo, hi = p.min(0) - self.radius, p.max(0) + self.radius
        zc = self.head[2]
        return (max(0, int(np.floor(lo[0])) - 1), min(self.nx, int(np.ceil(hi[0])) + 2),
                max(0, int(np.floor(lo[1])) - 1), min(self.ny, int(np.ceil(hi[1])) + 2),
                max(0, int(np.floor(zc - self.radius)) - 1), min(self.nz, int(np.ceil(zc + self.radius)) + 2))


def eel3d(seed: int = 4004, nx: int = 384, ny: int = 192, nz: int = 192, length: float = 160.0,
          radius: float = 8.0) -> Eel3D:
    """The precision-sweep eel: 10 capsules, length 160, radius 8 on 384x192x192,
    A ~ U[0.2, 0.6], f ~ U[0.5, 2] per 1000 steps, phi ~ U[0, pi/2] (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    e = Eel3D(nx=nx, ny=ny, nz=nz, amplitude=float(rng.uniform(0.2, 0.6)), frequency=float(rng.uniform(0.5, 2.0)),
              phase_lag=float(rng.uniform(0.0, np.pi / 2)), length=length, radius=radius)
    v = e.max_speed()
    if v > EEL_VMAX:
        e.time_scale = EEL_VMAX / v
    return e
