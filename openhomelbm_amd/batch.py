"""Batched D2Q9 engine: many resident lattices of uneven sizes advanced
together (hl_batch_* in include/homelbm.h).

The reference batches worlds of one size along a leading nworld launch
dimension (LBM_Solver, envs/lbm/lbm_solver.py:41-141). Here every
environment is an independent Lattice with its own extents, and the HIP
kernels walk (environment, strip) work-lists, so one launch per kernel kind
advances all of them. Prescribed kinematics come as per-step capsule and pose
tables; the mask is re-voxelized on the device at the start of each step.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .lattice import Lattice, _stream_ptr

MAX_SOLIDS = 16


class Batch:
    """A set of resident D2Q9 FP32 lattices stepped together.

    Args:
        lattices: the environments. Their moment buffers, masks and pose
            tables stay owned by the Lattice objects; the batch holds a copy of
            their descriptors, so change bodies through set_kinematics.
    """

    def __init__(self, lattices: Sequence[Lattice]):
        if not lattices:
            raise ValueError("a batch needs at least one lattice")
        for i, L in enumerate(lattices):
            if L.lattice != 9 or L.dtype != "f32":
                raise ValueError(f"lattice {i}: batches hold D2Q9 FP32 lattices")
            if L.nsolid > MAX_SOLIDS:
                raise ValueError(f"lattice {i}: at most {MAX_SOLIDS} solids per batched environment")
            if L.force is not None:
                raise ValueError(f"lattice {i}: batched environments have no body-force field")
        self.lib = N.load()
        self.lattices: List[Lattice] = list(lattices)
        self.n = len(self.lattices)
        self.device = self.lattices[0].device
        envs = (N.HLEnv * self.n)(*[L.env() for L in self.lattices])
        h = C.c_void_p()
        N.check(self.lib.hl_batch_create(envs, self.n, C.byref(h)), "hl_batch_create")
        self._h = h
        self.time = [0] * self.n
        self._tables = [None] * self.n
        self._active_key = None
        self._active_dev: Optional[torch.Tensor] = None
        self._graph = None  # (active key, nsteps) of the captured graph
        self._push_state()

    # ------------------------------------------------------------ state
    def _push_state(self):
        par = (C.c_int32 * self.n)(*[L.cur for L in self.lattices])
        tim = (C.c_int32 * self.n)(*self.time)
        N.check(self.lib.hl_batch_set_state(self._h, par, tim), "hl_batch_set_state")

    def pull_state(self):
        """Device parity / time of every environment (checks the host mirror)."""
        par = (C.c_int32 * self.n)()
        tim = (C.c_int32 * self.n)()
        N.check(self.lib.hl_batch_get_state(self._h, par, tim), "hl_batch_get_state")
        return list(par), list(tim)

    def set_time(self, i: int, t: int) -> None:
        self.time[i] = int(t)
        self._push_state()

    def set_strips(self, rows: int = 16, body_rows: int = 4) -> None:
        """Strip heights of the batched step (hl_batch_set_strips): rows away
        from the body boxes, body_rows where a column strip crosses one, both
        in [3, 64]. States do not depend on them; the force history agrees to
        summation-order rounding."""
        N.check(self.lib.hl_batch_set_strips(self._h, int(rows), int(body_rows)), "hl_batch_set_strips")
        self.strips = (int(rows), int(body_rows))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.hl_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sizes(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        N.check(self.lib.hl_batch_size(self._h, C.byref(a), C.byref(b), C.byref(c)), "hl_batch_size")
        return a.value, b.value, c.value

    @property
    def cells(self) -> int:
        return sum(L.ncell for L in self.lattices)

    # ----------------------------------------------------------- bodies
    def set_kinematics(self, i: int, caps: Optional[np.ndarray], poses: Optional[np.ndarray], box=None) -> None:
        """Per-step body tables of environment i, indexed by its time % T.

        caps: [T, ncap, 8] (ax, ay, az, bx, by, bz, radius, solid_id) or None
        (static mask already installed on the lattice); poses: [T, nsolid, 8]
        (px, py, cos, sin | last px, py, cos, sin) or None (the lattice's
        static poses). box: half-open cell box holding every capsule at every
        time (default: computed from the table)."""
        L = self.lattices[i]
        T = 0
        caps_dev = poses_dev = None
        ncap = 0
        if caps is not None:
            c = np.ascontiguousarray(caps, dtype=np.float32)
            if c.ndim != 3 or c.shape[2] != 8:
                raise ValueError("caps must be [T, ncap, 8]")
            if c.shape[1] and (c[..., 7].min() < 0 or c[..., 7].max() >= max(L.nsolid, 1) or c[..., 7].max() > 254):
                raise ValueError("capsule solid ids must be in [0, nsolid)")
            T, ncap = c.shape[0], c.shape[1]
            caps_dev = torch.from_numpy(c).to(self.device)
            if box is None:
                from .lattice import capsule_box
                box = capsule_box(c.reshape(-1, 8), (L.nx, L.ny, 1))
        if poses is not None:
            p = np.ascontiguousarray(poses, dtype=np.float32)
            if p.ndim != 3 or p.shape[1] != L.nsolid or p.shape[2] != 8:
                raise ValueError(f"poses must be [T, {L.nsolid}, 8]")
            if T and p.shape[0] != T:
                raise ValueError("caps and poses must have the same number of steps")
            T = p.shape[0]
            poses_dev = torch.from_numpy(p).to(self.device)
        if box is None:
            box = L.mask_box
        box = tuple(int(v) for v in box)
        if L.mask is None:
            L.mask = torch.zeros(L.ncell, dtype=torch.uint8, device=self.device)
        elif caps_dev is not None:
            L.mask.zero_()
        b = (C.c_int32 * 6)(*box)
        N.check(self.lib.hl_batch_set_bodies(
            self._h, i, L.mask.data_ptr(), b, caps_dev.data_ptr() if caps_dev is not None else None, ncap,
            poses_dev.data_ptr() if poses_dev is not None else None, T), "hl_batch_set_bodies")
        L.mask_box = box
        L.mask_changed()
        self._tables[i] = (caps_dev, poses_dev)
        self._graph = None

    # ------------------------------------------------------------- step
    def _active(self, active):
        if active is None:
            return None, None
        a = np.asarray(active, dtype=bool).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"active must have {self.n} entries")
        key = a.tobytes()
        if key != self._active_key:
            self._active_dev = torch.from_numpy(a.astype(np.uint8)).to(self.device)
            self._active_key = key
        return self._active_dev, a

    def _advance_mirror(self, a, nsteps):
        for k, L in enumerate(self.lattices):  # the voxelizer rewrote moving bodies' masks
            t = self._tables[k]
            if t is not None and t[0] is not None:
                L.mask_changed()
        for i, L in enumerate(self.lattices):
            if a is None or a[i]:
                L.cur ^= nsteps & 1
                self.time[i] += nsteps

    def history_buffer(self, nsteps: int) -> torch.Tensor:
        """Device FP64 [nsteps, n, 16, 6] force history."""
        return torch.zeros((nsteps, self.n, MAX_SOLIDS, 6), dtype=torch.float64, device=self.device)

    def _check_history(self, h, nsteps):
        if h is None:
            return None
        if (h.dtype != torch.float64 or not h.is_contiguous() or h.device != self.device
                or h.numel() < nsteps * self.n * MAX_SOLIDS * 6):
            raise ValueError("force_history must be a contiguous FP64 device tensor [nsteps, n, 16, 6]")
        return h.data_ptr()

    def step(self, nsteps: int = 1, active=None, force_history: Optional[torch.Tensor] = None,
             per_env: bool = False, stream=None) -> None:
        """Advance the (active) environments nsteps. per_env launches the
        same kernels one environment at a time (baseline schedule)."""
        act_dev, a = self._active(active)
        fh = self._check_history(force_history, nsteps)
        N.check(self.lib.hl_batch_step(self._h, act_dev.data_ptr() if act_dev is not None else None, int(nsteps),
                                       fh, 1 if per_env else 0, _stream_ptr(stream)), "hl_batch_step")
        self._advance_mirror(a, nsteps)

    def run(self, nsteps: int = 1, active=None, force_history: Optional[torch.Tensor] = None,
            per_env: bool = False, stream=None) -> None:
        """Advance the (active) environments nsteps in one launch of the
        persistent engine (per_env: one launch per environment); states equal
        step()'s bitwise."""
        act_dev, a = self._active(active)
        fh = self._check_history(force_history, nsteps)
        N.check(self.lib.hl_batch_run(self._h, act_dev.data_ptr() if act_dev is not None else None, int(nsteps),
                                      fh, 1 if per_env else 0, _stream_ptr(stream)), "hl_batch_run")
        self._advance_mirror(a, nsteps)

    def run_timed_out(self) -> bool:
        """Whether the last run() gave up a grid-barrier wait (synchronizes)."""
        v = C.c_int32(0)
        N.check(self.lib.hl_batch_run_status(self._h, C.byref(v)), "hl_batch_run_status")
        return bool(v.value)

    def capture(self, nsteps: int, active=None, force_history: Optional[torch.Tensor] = None) -> None:
        """Capture nsteps batched steps into a graph replayed by replay()."""
        act_dev, a = self._active(active)
        fh = self._check_history(force_history, nsteps)
        N.check(self.lib.hl_batch_capture(self._h, act_dev.data_ptr() if act_dev is not None else None, int(nsteps),
                                          fh), "hl_batch_capture")
        self._graph = (None if a is None else a.copy(), int(nsteps), force_history, act_dev)

    def replay(self, stream=None) -> None:
        if self._graph is None:
            raise RuntimeError("no captured graph: call capture() after the last set_kinematics()")
        N.check(self.lib.hl_batch_replay(self._h, _stream_ptr(stream)), "hl_batch_replay")
        self._advance_mirror(self._graph[0], self._graph[1])
