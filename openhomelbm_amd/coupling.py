"""Two-way coupling of lattices with a rigid-body engine, on the device.

The reference's coupled environments step MuJoCo-Warp and move data between
it and the solver with Warp kernels: body poses into the flows
(convert_and_update_solid_batch_2d, envs/lbm/lbm_fluid_env_func.py:72-105;
convert_and_update_solid_batch_3d, envs/lbm3d/lbm_fluid_env_3d_func.py:42-99)
and per-solid loads back into the engine (extract_forces_torques_batch,
lbm_fluid_env_func.py:51-69; extract_forces_torques_batch_3d and
extract_forces_torques_physical_3d, lbm_fluid_env_3d_func.py:102-185;
fill_xfrc_kernel / fill_xfrc_3d_kernel). Here those transfers are HIP kernels
(hl_couple_poses, hl_couple_forces) over device buffers. MuJoCo-Warp does not
run on this stack, so FreeBodies (hl_free_bodies_step) stands in for the
engine with free rigid bodies, and CoupledWorlds runs the whole loop: posed
capsules -> masks, poses -> pose tables, lattice step with its FP64 force
rows, rows -> engine loads, engine step; no host round trip per step.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .lattice import Lattice, _stream_ptr


def _dev(a, dtype, device) -> torch.Tensor:
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device).contiguous()


class BodyMap:
    """Engine body b <-> lattice solid solid_ids[b] with the mapping
    lattice = (engine - engine_origin[b]) * scale[b] + lattice_origin[b]
    (shared by every world, as the reference's mapping tables are)."""

    def __init__(self, solid_ids: Sequence[int], engine_origin=None, lattice_origin=None, scale=None,
                 body_ids: Optional[Sequence[int]] = None, device=None):
        n = len(solid_ids)
        self.device = torch.device(device or "cuda")
        self.nbody = n
        self.solid_ids = _dev(np.asarray(solid_ids, np.int32), torch.int32, self.device)
        eo = np.zeros((n, 3)) if engine_origin is None else np.asarray(engine_origin, np.float64).reshape(n, 3)
        lo = np.zeros((n, 3)) if lattice_origin is None else np.asarray(lattice_origin, np.float64).reshape(n, 3)
        sc = np.ones(n) if scale is None else np.asarray(scale, np.float64).reshape(n)
        self.engine_origin = _dev(eo, torch.float32, self.device)
        self.lattice_origin = _dev(lo, torch.float32, self.device)
        self.scale = _dev(sc, torch.float32, self.device)
        self.body_ids = None if body_ids is None else _dev(np.asarray(body_ids, np.int32), torch.int32, self.device)


def _tables(lattices: Sequence[Lattice]) -> torch.Tensor:
    ptr = np.array([lat.solids.data_ptr() for lat in lattices], np.uint64)
    return torch.from_numpy(ptr.view(np.int64)).to(lattices[0].device)


def poses_to_lattices(lattices: Sequence[Lattice], bmap: BodyMap, pos: torch.Tensor, quat: torch.Tensor,
                      initialized: torch.Tensor, pos_stride: int = 3, quat_stride: int = 4, tables=None,
                      stream=None) -> None:
    """Write the engine poses of every world (pos / quat device FP32, world
    major, body b of world w at (w * nbody + b) * stride) into the worlds'
    pose tables; initialized: device int32 [nworld, nsolid]."""
    lat0 = lattices[0]
    t = _tables(lattices) if tables is None else tables
    N.check(lat0.lib.hl_couple_poses(lat0.lattice, len(lattices), bmap.nbody, lat0.nsolid, pos.data_ptr(),
                                     pos_stride, quat.data_ptr(), quat_stride, bmap.solid_ids.data_ptr(),
                                     bmap.engine_origin.data_ptr(), bmap.lattice_origin.data_ptr(),
                                     bmap.scale.data_ptr(), t.data_ptr(), initialized.data_ptr(), _stream_ptr(stream)),
            "hl_couple_poses")


def forces_to_engine(lattice: int, rows: torch.Tensor, nworld: int, nsolid: int, bmap: BodyMap,
                     mode: int = N.HL_COUPLE_SCALED, force_conv: float = 1.0, torque_conv: float = 1.0,
                     world_stride: Optional[int] = None, xfrc: Optional[torch.Tensor] = None,
                     stream=None):
    """Engine loads (forces, torques: device FP32 [nworld, nbody, 3]) from the
    FP64 force rows (world w at rows + w * world_stride doubles); with xfrc
    [nworld, nbodies, 6] also the spatial forces at bmap.body_ids."""
    dev = rows.device
    fo = torch.empty((nworld, bmap.nbody, 3), dtype=torch.float32, device=dev)
    to = torch.empty_like(fo)
    ws = 6 * nsolid if world_stride is None else int(world_stride)
    lib = N.load()
    xp, xb, bp = None, 0, None
    if xfrc is not None:
        if bmap.body_ids is None:
            raise ValueError("xfrc needs body_ids in the BodyMap")
        xp, xb, bp = xfrc.data_ptr(), xfrc.shape[1], bmap.body_ids.data_ptr()
    N.check(lib.hl_couple_forces(lattice, mode, nworld, bmap.nbody, nsolid, rows.data_ptr(), ws,
                                 bmap.solid_ids.data_ptr(), bmap.scale.data_ptr(), float(force_conv),
                                 float(torque_conv), fo.data_ptr(), to.data_ptr(), xp, xb, bp, _stream_ptr(stream)),
            "hl_couple_forces")
    return fo, to


class FreeBodies:
    """Free rigid bodies of nworld worlds (the engine stand-in): state
    [nworld, nbody, 13] = position, quaternion (w, x, y, z), velocity,
    angular velocity (world frame); params [nbody, 4] = mass, principal
    inertia (body frame); semi-implicit Euler steps on the device."""

    def __init__(self, nworld: int, pos, mass, inertia, quat=None, vel=None, device=None):
        self.device = torch.device(device or "cuda")
        pos = np.asarray(pos, np.float64).reshape(nworld, -1, 3)
        self.nworld, self.nbody = pos.shape[0], pos.shape[1]
        s = np.zeros((self.nworld, self.nbody, 13))
        s[..., 0:3] = pos
        s[..., 3] = 1.0
        if quat is not None:
            s[..., 3:7] = np.asarray(quat, np.float64).reshape(self.nworld, self.nbody, 4)
        if vel is not None:
            s[..., 7:10] = np.asarray(vel, np.float64).reshape(self.nworld, self.nbody, 3)
        self.state = _dev(s, torch.float32, self.device)
        p = np.zeros((self.nbody, 4))
        p[:, 0] = np.asarray(mass, np.float64).reshape(self.nbody)
        p[:, 1:4] = np.asarray(inertia, np.float64).reshape(self.nbody, 3)
        self.params = _dev(p, torch.float32, self.device)

    def step(self, forces: torch.Tensor, torques: torch.Tensor, dt: float, gravity=(0.0, 0.0, 0.0), stream=None):
        g = [float(v) for v in gravity]
        N.check(N.load().hl_free_bodies_step(self.nworld, self.nbody, self.state.data_ptr(), self.params.data_ptr(),
                                            forces.data_ptr(), torques.data_ptr(), g[0], g[1], g[2], float(dt),
                                            _stream_ptr(stream)), "hl_free_bodies_step")

    def capsules(self, bmap: BodyMap, shapes: torch.Tensor, out: Optional[torch.Tensor] = None, stream=None):
        """Posed capsules [nworld, nbody, 8] in lattice coordinates."""
        if out is None:
            out = torch.empty((self.nworld, self.nbody, 8), dtype=torch.float32, device=self.device)
        N.check(N.load().hl_body_capsules(self.nworld, self.nbody, self.state.data_ptr(), 13,
                                         self.state.data_ptr() + 12, 13, shapes.data_ptr(),
                                         bmap.solid_ids.data_ptr(), bmap.engine_origin.data_ptr(),
                                         bmap.lattice_origin.data_ptr(), bmap.scale.data_ptr(), out.data_ptr(),
                                         _stream_ptr(stream)), "hl_body_capsules")
        return out


class CoupledWorlds:
    """nworld lattices, each with the same free bodies (capsule shapes
    [nbody, 7] in body coordinates), stepped two-way coupled on the device:
    per step, posed capsules re-voxelized into `box`, poses into the pose
    tables (the previous pose becomes 'last'), one lattice step with its
    force rows, rows -> body loads (physical mode with force_conv /
    torque_conv), one body step of dt."""

    def __init__(self, lattices: Sequence[Lattice], bodies: FreeBodies, bmap: BodyMap, shapes, box,
                 dt: float = 1.0, gravity=(0.0, 0.0, 0.0), force_conv: float = 1.0, torque_conv: float = 1.0):
        if len(lattices) != bodies.nworld:
            raise ValueError("one lattice per world")
        self.lattices, self.bodies, self.bmap = list(lattices), bodies, bmap
        dev = bodies.device
        self.shapes = _dev(np.asarray(shapes, np.float64).reshape(bmap.nbody, 7), torch.float32, dev)
        self.box = tuple(int(v) for v in box)
        self.dt, self.gravity = float(dt), tuple(float(v) for v in gravity)
        self.fc, self.tc = float(force_conv), float(torque_conv)
        lat0 = self.lattices[0]
        self.nsolid = lat0.nsolid
        self.rows = torch.zeros((bodies.nworld, 1, self.nsolid, 6), dtype=torch.float64, device=dev)
        self.init = torch.zeros((bodies.nworld, self.nsolid), dtype=torch.int32, device=dev)
        self.tables = _tables(self.lattices)
        self.caps = torch.empty((bodies.nworld, bmap.nbody, 8), dtype=torch.float32, device=dev)

    def step(self, nsteps: int = 1, stream=None) -> None:
        b = self.bodies
        for _ in range(nsteps):
            b.capsules(self.bmap, self.shapes, out=self.caps, stream=stream)
            for w, lat in enumerate(self.lattices):
                lat.voxelize_device(self.caps[w], self.box, stream=stream)
            poses_to_lattices(self.lattices, self.bmap, b.state, b.state[..., 3:], self.init, 13, 13,
                              tables=self.tables, stream=stream)
            for w, lat in enumerate(self.lattices):
                lat.step(1, N.HL_KERNEL_AUTO, force_history=self.rows[w], stream=stream)
            f, t = forces_to_engine(self.lattices[0].lattice, self.rows, b.nworld, self.nsolid, self.bmap,
                                    N.HL_COUPLE_PHYSICAL, self.fc, self.tc, stream=stream)
            b.step(f, t, self.dt, self.gravity, stream=stream)
